"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  See oracle/oracle.h for
what each function restates from the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp, i32, f32p = C.c_void_p, C.c_int32, C.c_void_p
    L.orc_scene_create.restype = vp
    L.orc_scene_create.argtypes = [i32, vp, vp, vp, i32, vp, vp, i32, vp, i32, vp, vp, vp]
    L.orc_hit_material.restype = None
    L.orc_hit_material.argtypes = [vp, vp, i32, vp]
    L.orc_image_texel.restype = None
    L.orc_image_texel.argtypes = [vp, i32, i32, i32, vp]
    L.orc_set_corrected.restype = None
    L.orc_set_corrected.argtypes = [vp, i32]
    L.orc_render_lt.restype = i32
    L.orc_render_lt.argtypes = [vp, i32, i32, i32, i32, i32, C.c_uint64, vp, i32, vp]
    L.orc_scene_set_point_lights.restype = None
    L.orc_scene_set_point_lights.argtypes = [vp, i32, vp]
    L.orc_scene_set_textures.restype = None
    L.orc_scene_set_textures.argtypes = [vp, vp, i32, vp, i32, vp, vp]
    L.orc_scene_destroy.argtypes = [vp]
    L.orc_scene_destroy.restype = None
    L.orc_bvh_nodes.argtypes = [vp]
    L.orc_bvh_rank.argtypes = [vp, vp]
    L.orc_bvh_rank.restype = None
    L.orc_trace_closest.argtypes = [vp, vp, i32, i32, vp, vp, i32]
    L.orc_trace_closest.restype = None
    L.orc_trace_any.argtypes = [vp, vp, i32, i32, vp, i32]
    L.orc_trace_any.restype = None
    L.orc_camera_ray.argtypes = [vp, C.c_float, C.c_float, vp, vp]
    L.orc_camera_ray.restype = None
    L.orc_render.argtypes = [vp, i32, i32, i32, i32, i32, C.c_uint64, i32, i32, i32, i32, vp, vp, vp, i32, vp]
    L.orc_path.argtypes = [vp, i32, i32, i32, i32, C.c_uint64, i32, i32, vp]
    L.orc_set_integrator.argtypes = [vp, i32]
    L.orc_set_view.argtypes = [vp, C.c_uint32]
    L.orc_set_view.restype = None
    L.orc_ppg_render.argtypes = [vp, i32, i32, i32, i32, C.c_uint64, C.c_uint32] + [vp] * 5 + [i32, vp, vp, vp, i32]
    L.orc_path.restype = None
    L.orc_diffuse_sample.argtypes = [vp, vp, C.c_float, C.c_float, vp, vp]
    L.orc_diffuse_sample.restype = None
    L.orc_glossy_sample.argtypes = [vp, vp, vp, C.c_float, vp, C.c_float, C.c_float, vp, vp, vp]
    L.orc_glossy_sample.restype = None
    L.orc_glossy_eval.argtypes = [vp, vp, vp, C.c_float, vp, vp]
    L.orc_glossy_eval.restype = C.c_float
    L.orc_glossy_pdf.argtypes = [vp, vp, vp, C.c_float, vp, vp]
    L.orc_glossy_pdf.restype = C.c_float
    L.orc_ggx_d.argtypes = [C.c_float, vp]
    L.orc_ggx_d.restype = C.c_float
    L.orc_ggx_pdf.argtypes = [C.c_float, vp, vp]
    L.orc_ggx_pdf.restype = C.c_float
    L.orc_ggx_sample.argtypes = [C.c_float, vp, C.c_float, C.c_float, vp]
    L.orc_ggx_sample.restype = None
    L.orc_triangle_sample.argtypes = [vp, i32, C.c_float, C.c_float, vp, vp, vp]
    L.orc_triangle_sample.restype = None
    L.orc_pcg_floats.argtypes = [C.c_uint64, C.c_uint64, i32, vp]
    L.orc_pcg_floats.restype = None
    L.orc_ggx_lambda.argtypes = [C.c_float, vp]
    L.orc_ggx_lambda.restype = C.c_float
    L.orc_ggx_g1.argtypes = [C.c_float, vp]
    L.orc_ggx_g1.restype = C.c_float
    L.orc_dielectric_sample.argtypes = [vp, vp, C.c_float, vp]
    L.orc_dielectric_sample.restype = None
    L.orc_russian_roulette.argtypes = [vp, i32]
    L.orc_russian_roulette.restype = C.c_float
    for f in ("orc_sd_lookup", "orc_sd_pdf", "orc_sd_sample"):
        getattr(L, f).restype = None
    L.orc_sd_lookup.argtypes = [vp] * 5 + [i32, vp, vp]
    L.orc_sd_pdf.argtypes = [vp] * 5 + [i32, vp, vp, vp]
    L.orc_sd_sample.argtypes = [vp] * 5 + [i32, vp, vp, vp, vp]
    L.orc_ggx_sample_batch.argtypes = [C.c_float, vp, i32, vp, vp]
    L.orc_ggx_sample_batch.restype = None
    L.orc_mt_floats.argtypes = [C.c_uint32, i32, vp]
    L.orc_mt_floats.restype = None
    L.orc_mlp_num_params.restype = i32
    L.orc_mlp_set_width.argtypes = [i32]
    L.orc_mlp_set_width.restype = None
    L.orc_mlp_forward.argtypes = [vp, vp, i32, i32, vp]
    L.orc_mlp_forward.restype = None
    L.orc_mlp_mixture_pdf_batch.argtypes = [vp, vp, i32, vp]
    L.orc_mlp_mixture_pdf_batch.restype = None
    L.orc_mlp_loss_grad.argtypes = [vp, vp, i32, i32, vp]
    L.orc_mlp_loss_grad.restype = C.c_double
    L.orc_mlp_adam.argtypes = [vp, vp, vp, vp, i32, C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_mlp_adam.restype = None
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _v3(x):
    return np.ascontiguousarray(np.asarray(x, np.float32).reshape(3))


class OracleScene:
    """A scene (cpt.scenes.Scene) loaded into the oracle."""

    def __init__(self, scene, camera=None):
        a = scene.c_arrays() if hasattr(scene, "c_arrays") else scene
        self.arrays = a
        cam = np.ascontiguousarray(a["camera"] if camera is None else camera, np.float32)
        self._cam = cam
        L = lib()
        self.h = L.orc_scene_create(
            a["positions"].shape[0], _p(a["positions"]), _p(a["normals"]), _p(a["bitangents"]),
            a["indices"].shape[0], _p(a["indices"]), _p(a["tri_material"]),
            a["materials"].shape[0], _p(a["materials"]), a["light_tris"].shape[0], _p(a["light_tris"]),
            _p(cam), _p(np.ascontiguousarray(a["env"], np.float32)))
        if "textures" in a and len(a["textures"]):
            L.orc_scene_set_textures(self.h, _p(a["uvs"]), a["textures"].shape[0], _p(a["textures"]),
                                     a["texels"].size, _p(a["texels"]), _p(a["material_textures"]))
        pl = a.get("point_lights") if hasattr(a, "get") else None
        if pl is not None and len(pl):
            L.orc_scene_set_point_lights(self.h, pl.shape[0], _p(pl))
        self.n_triangles = a["indices"].shape[0]

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bvh_nodes(self):
        return lib().orc_bvh_nodes(self.h)

    def rank(self):
        r = np.zeros(self.n_triangles, np.int32)
        lib().orc_bvh_rank(self.h, _p(r))
        return r

    def trace_closest(self, rays, brute=False, threads=0):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        lib().orc_trace_closest(self.h, _p(rays), n, 1 if brute else 0, _p(tri), _p(t), threads)
        return tri, t

    def hit_material(self, rays):
        """Closest-hit material after textures: (n, 10) = type (-1 miss), albedo,
        roughness, normal-map texel, u, v."""
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((rays.shape[0], 10), np.float32)
        lib().orc_hit_material(self.h, _p(rays), rays.shape[0], _p(out))
        return out

    def image_texel(self, img, i, j):
        out = np.zeros(3, np.float32)
        lib().orc_image_texel(self.h, img, i, j, _p(out))
        return out[:self.arrays["textures"][img][2]]

    def trace_any(self, rays, brute=False, threads=0):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint8)
        lib().orc_trace_any(self.h, _p(rays), n, 1 if brute else 0, _p(occ), threads)
        return occ

    def camera_ray(self, cx, cy):
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.float32)
        lib().orc_camera_ray(self.h, cx, cy, _p(o), _p(d))
        return o, d

    def render(self, W, H, spp_axis, max_depth=32, sampler="pcg", seed=1, cols=None, rows=None,
               threads=0, sumsq=False, integrator=0, corrected=False, view=0):
        """Returns (sums (H,W,3) float64, film (H,W,3) float32, counts [paths, ext, shadow])
        and, with sumsq=True, the per-pixel sums of squares as a fourth element.
        sampler "pcg" = the GPU's per-path stream; "mt" = the reference's per-column
        mt19937 (seed 0 reproduces the reference's seeding).  integrator 0 =
        PathtracerMIS, 1 = Pathtracer (NEE without MIS), 2 = PurePt."""
        lib().orc_set_integrator(self.h, integrator)
        lib().orc_set_corrected(self.h, int(corrected))
        lib().orc_set_view(self.h, view)
        sums = np.zeros((H, W, 3), np.float64)
        sq = np.zeros((H, W, 3), np.float64)
        film = np.zeros((H, W, 3), np.float32)
        counts = np.zeros(3, np.int64)
        c0, c1 = cols if cols else (0, W)
        r0, r1 = rows if rows else (0, H)
        lib().orc_render(self.h, W, H, spp_axis, max_depth, 0 if sampler == "pcg" else 1, seed,
                         c0, c1, r0, r1, _p(sums), _p(sq), _p(film), threads, _p(counts))
        return (sums, film, counts, sq) if sumsq else (sums, film, counts)

    def render_lt(self, W, H, spp_axis, max_depth=32, sampler="pcg", seed=1, threads=0, corrected=False):
        """Lighttracer: returns (film sums (H,W,3) float64 = spp x the image, counts)."""
        lib().orc_set_corrected(self.h, int(corrected))
        sums = np.zeros((H, W, 3), np.float64)
        counts = np.zeros(3, np.int64)
        lib().orc_render_lt(self.h, W, H, spp_axis, max_depth, 0 if sampler == "pcg" else 1, seed, _p(sums),
                            threads, _p(counts))
        return sums, counts

    def ppg_render(self, W, H, spp, max_depth, seed, view, tree, threads=0):
        """One PPG iteration with the fixed SD-tree `tree` (a dict as returned by
        cpt.Context.ppg_export); returns (pixel sums (H,W,3), splat sums, splat counts)."""
        sums = np.zeros((H, W, 3), np.float64)
        ns = len(tree["sums"])
        ssum = np.zeros(ns, np.float64)
        scnt = np.zeros(ns, np.int64)
        t = {k: np.ascontiguousarray(v) for k, v in tree.items()}
        lib().orc_set_corrected(self.h, 0)
        lib().orc_ppg_render(self.h, W, H, spp, max_depth, seed, view, _p(t["snode"]), _p(t["dinfo"]), _p(t["quad"]),
                             _p(t["rect"]), _p(t["bounds"]), ns, _p(sums), _p(ssum), _p(scnt), threads)
        return sums, ssum, scnt

    def path(self, W, H, spp_axis, max_depth, seed, pixel, sample, integrator=0):
        out = np.zeros(3, np.float32)
        lib().orc_set_integrator(self.h, integrator)
        lib().orc_set_corrected(self.h, 0)
        lib().orc_path(self.h, W, H, spp_axis, max_depth, seed, pixel, sample, _p(out))
        return out

    def triangle_sample(self, tri, u1, u2):
        pos = np.zeros(3, np.float32)
        nrm = np.zeros(3, np.float32)
        area = np.zeros(1, np.float32)
        lib().orc_triangle_sample(self.h, tri, u1, u2, _p(pos), _p(nrm), _p(area))
        return pos, nrm, float(area[0])


def diffuse_sample(n, wo, u1, u2):
    wi = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().orc_diffuse_sample(_p(_v3(n)), _p(_v3(wo)), u1, u2, _p(wi), _p(pdf))
    return wi, float(pdf[0])


def glossy_sample(t, b, n, rough, wo, u1, u2):
    wi = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    delta = np.zeros(1, np.int32)
    lib().orc_glossy_sample(_p(_v3(t)), _p(_v3(b)), _p(_v3(n)), rough, _p(_v3(wo)), u1, u2, _p(wi), _p(pdf), _p(delta))
    return wi, float(pdf[0]), bool(delta[0])


def glossy_eval(t, b, n, rough, wo, wi):
    return lib().orc_glossy_eval(_p(_v3(t)), _p(_v3(b)), _p(_v3(n)), rough, _p(_v3(wo)), _p(_v3(wi)))


def glossy_pdf(t, b, n, rough, wo, wi):
    return lib().orc_glossy_pdf(_p(_v3(t)), _p(_v3(b)), _p(_v3(n)), rough, _p(_v3(wo)), _p(_v3(wi)))


def ggx_lambda(alpha, w):
    return lib().orc_ggx_lambda(alpha, _p(_v3(w)))


def ggx_g1(alpha, w):
    return lib().orc_ggx_g1(alpha, _p(_v3(w)))


def dielectric_sample(n, wo, ior):
    wi = np.zeros(3, np.float32)
    lib().orc_dielectric_sample(_p(_v3(n)), _p(_v3(wo)), ior, _p(wi))
    return wi


def russian_roulette(c, corrected=False):
    return lib().orc_russian_roulette(_p(_v3(c)), int(corrected))


def ggx_sample_batch(alpha, wo, u):
    """Visible-normal samples for an (n, 2) array of uniforms."""
    u = np.ascontiguousarray(u, np.float32)
    out = np.zeros((u.shape[0], 3), np.float32)
    lib().orc_ggx_sample_batch(alpha, _p(_v3(wo)), u.shape[0], _p(u), _p(out))
    return out


def _sd_arrays(tree):
    t = {k: np.ascontiguousarray(tree[k], np.int32 if k in ("snode", "dinfo", "quad") else np.float32)
         for k in ("snode", "dinfo", "quad", "rect", "bounds")}
    return t, [_p(t[k]) for k in ("snode", "dinfo", "quad", "rect", "bounds")]


def sd_lookup(tree, pts):
    """Directional tree index of each point (flattened SD-tree as cpt_ppg_export gives it)."""
    keep, a = _sd_arrays(tree)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.zeros(p.shape[0], np.int32)
    lib().orc_sd_lookup(*a, p.shape[0], _p(p), _p(out))
    return out


def sd_pdf(tree, pts, dirs):
    keep, a = _sd_arrays(tree)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.zeros(p.shape[0], np.float32)
    lib().orc_sd_pdf(*a, p.shape[0], _p(p), _p(d), _p(out))
    return out


def sd_sample(tree, pts, u3):
    keep, a = _sd_arrays(tree)
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    u = np.ascontiguousarray(u3, np.float32).reshape(-1, 3)
    dirs = np.zeros((p.shape[0], 3), np.float32)
    pdf = np.zeros(p.shape[0], np.float32)
    lib().orc_sd_sample(*a, p.shape[0], _p(p), _p(u), _p(dirs), _p(pdf))
    return dirs, pdf


def ggx_d(alpha, wh):
    return lib().orc_ggx_d(alpha, _p(_v3(wh)))


def ggx_pdf(alpha, wo, wh):
    return lib().orc_ggx_pdf(alpha, _p(_v3(wo)), _p(_v3(wh)))


def ggx_sample(alpha, wo, u1, u2):
    out = np.zeros(3, np.float32)
    lib().orc_ggx_sample(alpha, _p(_v3(wo)), u1, u2, _p(out))
    return out


def pcg_floats(seed, path_id, n):
    out = np.zeros(n, np.float32)
    lib().orc_pcg_floats(seed, path_id, n, _p(out))
    return out


def mt_floats(seed, n):
    out = np.zeros(n, np.float32)
    lib().orc_mt_floats(seed, n, _p(out))
    return out


# ---------------------------------------------------------------- guide MLP oracle
def _mlp_width(width):
    if width not in (64, 128):
        raise ValueError("guide network width is 64 or 128")
    lib().orc_mlp_set_width(width)
    return 48 if width == 64 else 80


def mlp_num_params(width=64):
    _mlp_width(width)
    return lib().orc_mlp_num_params()


def mlp_forward(params, inputs, bf16=True, width=64):
    nout = _mlp_width(width)
    P = np.ascontiguousarray(params, np.float32)
    assert P.size == lib().orc_mlp_num_params()
    X = np.ascontiguousarray(inputs, np.float32).reshape(-1, 7)
    out = np.zeros((X.shape[0], nout), np.float64)
    lib().orc_mlp_forward(_p(P), _p(X), X.shape[0], 1 if bf16 else 0, _p(out))
    return out


def mlp_mixture_pdf(outputs, dirs, width=64):
    nout = _mlp_width(width)
    o = np.ascontiguousarray(outputs, np.float64).reshape(nout)
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    out = np.zeros(d.shape[0], np.float64)
    lib().orc_mlp_mixture_pdf_batch(_p(o), _p(d), d.shape[0], _p(out))
    return out


def mlp_loss_grad(params, samples, bf16=True, width=64):
    _mlp_width(width)
    P = np.ascontiguousarray(params, np.float32)
    assert P.size == lib().orc_mlp_num_params()
    S = np.ascontiguousarray(samples, np.float32).reshape(-1, 11)
    g = np.zeros(P.size, np.float64)
    loss = lib().orc_mlp_loss_grad(_p(P), _p(S), S.shape[0], 1 if bf16 else 0, _p(g))
    return loss, g


def mlp_adam(params, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    P = np.array(params, np.float32)
    m = np.array(m, np.float64)
    v = np.array(v, np.float64)
    g = np.ascontiguousarray(grad, np.float64)
    _mlp_width(64 if P.size == mlp_num_params(64) else 128)
    assert P.size == m.size == v.size == g.size == lib().orc_mlp_num_params()
    lib().orc_mlp_adam(_p(P), _p(m), _p(v), _p(g), step, lr, b1, b2, eps)
    return P, m, v
