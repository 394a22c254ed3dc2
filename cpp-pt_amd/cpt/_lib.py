"""ctypes binding of libcpt.so (the C ABI declared in include/cpt.h).

The library is loaded from the package directory (cpp-pt_amd/libcpt.so).  There
is no fallback: if the library is missing or a call fails, an exception is
raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libcpt.so")

CPT_OK = 0
CPT_E_INVALID = -1
CPT_E_HIP = -2
CPT_E_NOSCENE = -3
CPT_E_NODEVICE = -4
CPT_E_LIMIT = -5

CPT_SCENE_NO_TIE_RANK = 1
CPT_SCENE_GPU_BVH = 2
CPT_RENDER_NO_GRAPH = 1
CPT_RENDER_KERNEL_TIMING = 2
CPT_RENDER_COUNT = 4
CPT_RENDER_ACCUMULATE = 8
CPT_RENDER_STREAM = 16
CPT_RENDER_GUIDED = 32
CPT_RENDER_GUIDE_TRAIN = 64
CPT_RENDER_NEE_ONLY = 128
CPT_RENDER_PURE_PT = 256
CPT_RENDER_PPG = 512
CPT_RENDER_PPG_STEP = 1024
CPT_RENDER_LIGHT_TRACE = 2048
CPT_RENDER_CORRECTED = 4096

# every function include/cpt.h declares
ABI_FUNCTIONS = (
    "cpt_version", "cpt_create", "cpt_destroy", "cpt_last_error", "cpt_upload_scene",
    "cpt_bvh_info", "cpt_tie_rank", "cpt_trace_closest", "cpt_trace_any",
    "cpt_trace_closest_device", "cpt_render", "cpt_read_film", "cpt_render_views",
    "cpt_reference_tie_rank", "cpt_guide_num_params", "cpt_guide_shape", "cpt_guide_init", "cpt_guide_get_params",
    "cpt_guide_set_params", "cpt_guide_forward", "cpt_guide_train_step", "cpt_autotune",
    "cpt_ppg_reset", "cpt_ppg_update", "cpt_ppg_export", "cpt_debug_fail_alloc",
    "cpt_trace_closest_device_noalloc", "cpt_scene_file_load", "cpt_scene_file_free",
    "cpt_guide_mixtures", "cpt_debug_guide_mixtures", "cpt_debug_quantize_node",
)


class GuideParams(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("seed", C.c_uint64), ("guide_fraction", C.c_float), ("train_paths", C.c_int32),
                ("train_steps", C.c_int32), ("guide_depth", C.c_int32), ("guide_target", C.c_int32),
                ("infer_fp8", C.c_int32), ("width", C.c_int32), ("fraction_ramp", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int32), ("positions", C.c_void_p), ("normals", C.c_void_p),
        ("bitangents", C.c_void_p), ("n_triangles", C.c_int32), ("indices", C.c_void_p),
        ("tri_material", C.c_void_p), ("n_materials", C.c_int32), ("materials", C.c_void_p),
        ("n_lights", C.c_int32), ("light_triangles", C.c_void_p), ("camera", C.c_float * 16),
        ("env", C.c_float * 3), ("flags", C.c_int32), ("uvs", C.c_void_p), ("n_textures", C.c_int32),
        ("textures", C.c_void_p), ("texels", C.c_void_p), ("material_textures", C.c_void_p),
        ("n_point_lights", C.c_int32), ("point_lights", C.c_void_p),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp_axis", C.c_int32),
        ("max_depth", C.c_int32), ("seed", C.c_uint64), ("flags", C.c_int32),
        ("max_paths", C.c_int32), ("block_size", C.c_int32), ("grid_blocks", C.c_int32),
        ("view", C.c_int32), ("shade_block", C.c_int32), ("refill", C.c_int32), ("reserved", C.c_int32),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("seconds", C.c_double), ("paths", C.c_int64), ("extension_rays", C.c_int64),
        ("shadow_rays", C.c_int64), ("node_visits", C.c_int64), ("triangle_tests", C.c_int64),
        ("shadow_node_visits", C.c_int64), ("shadow_triangle_tests", C.c_int64),
        ("ms_extend", C.c_double), ("ms_shade", C.c_double), ("ms_shadow", C.c_double),
        ("ms_other", C.c_double), ("n_extend", C.c_int64), ("n_shade", C.c_int64),
        ("n_shadow", C.c_int64), ("passes", C.c_int32), ("paths_per_pass", C.c_int32),
        ("block_size", C.c_int32), ("grid_blocks", C.c_int32), ("guide_samples", C.c_int64),
        ("guide_steps", C.c_int32), ("graph_captured", C.c_int32), ("ms_guide", C.c_double), ("n_guide", C.c_int64),
        ("ms_train", C.c_double), ("n_train", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32), ("grid_blocks", C.c_int32), ("shade_block", C.c_int32),
        ("refill", C.c_int32), ("trials", C.c_int32), ("reserved0", C.c_int32),
        ("samples_per_s", C.c_double), ("seconds", C.c_double), ("reserved", C.c_int32 * 4),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libcpt.so once; raise when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libcpt.so not found at {path}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    lib.cpt_version.restype = C.c_int
    lib.cpt_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    lib.cpt_destroy.argtypes = [C.c_void_p]
    lib.cpt_destroy.restype = None
    lib.cpt_last_error.argtypes = [C.c_void_p]
    lib.cpt_last_error.restype = C.c_char_p
    lib.cpt_debug_fail_alloc.argtypes = [C.c_void_p, C.c_int32]
    lib.cpt_debug_quantize_node.argtypes = [C.c_void_p, C.c_void_p]
    lib.cpt_upload_scene.argtypes = [C.c_void_p, C.POINTER(SceneDesc)]
    lib.cpt_bvh_info.argtypes = [C.c_void_p, C.c_void_p]
    lib.cpt_tie_rank.argtypes = [C.c_void_p, C.c_void_p]
    lib.cpt_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    lib.cpt_trace_any.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    lib.cpt_trace_closest_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    lib.cpt_trace_closest_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
    lib.cpt_render.argtypes = [C.c_void_p, C.POINTER(RenderParams), C.POINTER(RenderStats)]
    lib.cpt_autotune.argtypes = [C.c_void_p, C.POINTER(RenderParams), C.POINTER(TuneResult)]
    lib.cpt_ppg_reset.argtypes = [C.c_void_p]
    lib.cpt_ppg_update.argtypes = [C.c_void_p]
    lib.cpt_ppg_export.argtypes = [C.c_void_p] + [C.c_void_p] * 8
    lib.cpt_read_film.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.cpt_render_views.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                     C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p,
                                     C.POINTER(RenderStats)]
    lib.cpt_reference_tie_rank.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    lib.cpt_scene_file_load.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                        C.POINTER(SceneDesc), C.c_char_p, C.c_int32]
    lib.cpt_scene_file_free.argtypes = [C.c_void_p]
    lib.cpt_scene_file_free.restype = None
    lib.cpt_guide_num_params.restype = C.c_int
    lib.cpt_guide_init.argtypes = [C.c_void_p, C.POINTER(GuideParams)]
    lib.cpt_guide_shape.argtypes = [C.c_void_p, C.c_void_p]
    lib.cpt_guide_get_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.cpt_guide_set_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.cpt_guide_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    lib.cpt_guide_mixtures.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    lib.cpt_debug_guide_mixtures.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    lib.cpt_guide_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    _lib = lib
    return lib


def reference_tie_rank(positions, indices):
    """Host-only: the reference BVH test order of every triangle."""
    lib = load_library()
    P = np.ascontiguousarray(positions, np.float32)
    I = np.ascontiguousarray(indices, np.int32)
    out = np.zeros(I.shape[0], np.int32)
    rc = lib.cpt_reference_tie_rank(_ptr(P), _ptr(I), I.shape[0], _ptr(out))
    if rc != CPT_OK:
        raise CptError(rc, "cpt_reference_tie_rank")
    return out


def load_scene_file(path, res_x, res_y):
    """Host-only: parse a scene file through cpt_scene_file_load and return copies of
    the flattened arrays, keyed like scenes.Scene.c_arrays()."""
    lib = load_library()
    handle = C.c_void_p()
    desc = SceneDesc()
    err = C.create_string_buffer(512)
    rc = lib.cpt_scene_file_load(os.fsencode(path), int(res_x), int(res_y), C.byref(handle), C.byref(desc), err, 512)
    if rc != CPT_OK:
        raise CptError(rc, err.value.decode(errors="replace"))
    try:
        def arr(ptr, dtype, n, cols):
            if n == 0 or not ptr:
                return np.zeros((0, cols), dtype)
            ct = C.c_float if dtype == np.float32 else C.c_int32
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n * cols,))
            return a.copy().reshape(n, cols)
        nv, nt, nm = desc.n_vertices, desc.n_triangles, desc.n_materials
        out = dict(
            positions=arr(desc.positions, np.float32, nv, 3), normals=arr(desc.normals, np.float32, nv, 3),
            bitangents=arr(desc.bitangents, np.float32, nv, 3), indices=arr(desc.indices, np.int32, nt, 3),
            tri_material=arr(desc.tri_material, np.int32, nt, 1).reshape(-1),
            materials=arr(desc.materials, np.float32, nm, 16),
            light_tris=arr(desc.light_triangles, np.int32, desc.n_lights, 1).reshape(-1),
            point_lights=arr(desc.point_lights, np.float32, desc.n_point_lights, 6),
            camera=np.array(desc.camera[:], np.float32), env=np.array(desc.env[:], np.float32),
            uvs=arr(desc.uvs, np.float32, nv, 2), textures=arr(desc.textures, np.int32, desc.n_textures, 4),
            material_textures=arr(desc.material_textures, np.int32, nm, 4),
        )
        n_texels = max((int(o) + int(w) * int(h) * int(ch) for w, h, ch, o in out["textures"]), default=0)
        out["texels"] = arr(desc.texels, np.float32, n_texels, 1).reshape(-1)
        return out
    finally:
        lib.cpt_scene_file_free(handle)


class CptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libcpt error {code}: {msg}")
        self.code = code


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Context:
    """One libcpt context (one HIP stream on one gfx950 device)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.cpt_create(C.byref(h), device)
        if rc != CPT_OK:
            raise CptError(rc, f"cpt_create(device={device}) failed")
        self.h = h
        self._keep = None
        self.n_triangles = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.cpt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != CPT_OK:
            raise CptError(rc, f"{what}: {self.lib.cpt_last_error(self.h).decode()}")

    def debug_fail_alloc(self, after):
        """Make the device allocation after the next `after` fail (tests; -1: off)."""
        self._check(self.lib.cpt_debug_fail_alloc(self.h, after), "cpt_debug_fail_alloc")

    def upload(self, scene, tie_rank: bool = True, camera=None, gpu_bvh: bool = False):
        """Upload a cpt.scenes.Scene; `camera` (16 floats) overrides its camera;
        `gpu_bvh` builds the BVH on the device."""
        arrs = scene.c_arrays()
        d = SceneDesc()
        d.n_vertices = arrs["positions"].shape[0]
        d.positions = _ptr(arrs["positions"])
        d.normals = _ptr(arrs["normals"])
        d.bitangents = _ptr(arrs["bitangents"])
        d.n_triangles = arrs["indices"].shape[0]
        d.indices = _ptr(arrs["indices"])
        d.tri_material = _ptr(arrs["tri_material"])
        d.n_materials = arrs["materials"].shape[0]
        d.materials = _ptr(arrs["materials"])
        d.n_lights = arrs["light_tris"].shape[0]
        d.light_triangles = _ptr(arrs["light_tris"])
        cam = arrs["camera"] if camera is None else np.asarray(camera, np.float32).reshape(16)
        for k in range(16):
            d.camera[k] = float(cam[k])
        for k in range(3):
            d.env[k] = float(arrs["env"][k])
        d.flags = (0 if tie_rank else CPT_SCENE_NO_TIE_RANK) | (CPT_SCENE_GPU_BVH if gpu_bvh else 0)
        d.uvs = _ptr(arrs["uvs"])
        d.n_textures = arrs["textures"].shape[0]
        d.textures = _ptr(arrs["textures"]) if d.n_textures else None
        d.texels = _ptr(arrs["texels"]) if d.n_textures else None
        d.material_textures = _ptr(arrs["material_textures"])
        pl = arrs.get("point_lights")
        d.n_point_lights = 0 if pl is None else pl.shape[0]
        d.point_lights = _ptr(pl) if d.n_point_lights else None
        self._check(self.lib.cpt_upload_scene(self.h, C.byref(d)), "cpt_upload_scene")
        self.n_triangles = d.n_triangles

    def bvh_info(self):
        out = np.zeros(6, np.int64)
        self._check(self.lib.cpt_bvh_info(self.h, _ptr(out)), "cpt_bvh_info")
        return dict(nodes=int(out[0]), leaves=int(out[1]), depth=int(out[2]), triangles=int(out[3]),
                    build_ms=out[4] / 1000.0, gpu_built=bool(out[5]))

    def tie_rank(self):
        out = np.zeros(self.n_triangles, np.int32)
        self._check(self.lib.cpt_tie_rank(self.h, _ptr(out)), "cpt_tie_rank")
        return out

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        self._check(self.lib.cpt_trace_closest(self.h, _ptr(rays), n, _ptr(tri), _ptr(t)), "cpt_trace_closest")
        return tri, t

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint8)
        self._check(self.lib.cpt_trace_any(self.h, _ptr(rays), n, _ptr(occ)), "cpt_trace_any")
        return occ

    def render(self, width, height, spp_axis, max_depth=32, seed=1, flags=0, max_paths=0,
               block_size=0, grid_blocks=0, view=0, shade_block=0, refill=0):
        p = RenderParams()
        p.width, p.height, p.spp_axis, p.max_depth = width, height, spp_axis, max_depth
        p.seed, p.flags, p.max_paths = seed, flags, max_paths
        p.block_size, p.grid_blocks, p.view = block_size, grid_blocks, view
        p.shade_block = shade_block
        p.refill = refill
        st = RenderStats()
        self._check(self.lib.cpt_render(self.h, C.byref(p), C.byref(st)), "cpt_render")
        self._last = (width, height)
        return st.as_dict()

    def autotune(self, width, height, max_depth=32, seed=1, flags=0, max_paths=0, view=0):
        """Search launch configurations for this scene/size/depth (cpt_autotune);
        later renders with those parameters left 0 use the winner."""
        p = RenderParams()
        p.width, p.height, p.spp_axis, p.max_depth = width, height, 1, max_depth
        p.seed, p.flags, p.max_paths, p.view = seed, flags, max_paths, view
        r = TuneResult()
        self._check(self.lib.cpt_autotune(self.h, C.byref(p), C.byref(r)), "cpt_autotune")
        self._last = (width, height)
        return r.as_dict()

    def ppg_reset(self):
        self._check(self.lib.cpt_ppg_reset(self.h), "cpt_ppg_reset")

    def ppg_update(self):
        self._check(self.lib.cpt_ppg_update(self.h), "cpt_ppg_update")

    def ppg_export(self):
        """The flattened SD-tree the next PPG step uses and the current splats."""
        sizes = np.zeros(5, np.int32)
        nul = C.c_void_p(0)
        self._check(self.lib.cpt_ppg_export(self.h, _ptr(sizes), nul, nul, nul, nul, nul, nul, nul), "cpt_ppg_export")
        t = dict(snode=np.zeros(sizes[0], np.int32), dinfo=np.zeros((sizes[1], 4), np.int32),
                 quad=np.zeros(sizes[2], np.int32), rect=np.zeros((sizes[3], 4), np.float32),
                 sums=np.zeros(sizes[4], np.float32), counts=np.zeros(sizes[4], np.uint32),
                 bounds=np.zeros(6, np.float32))
        self._check(self.lib.cpt_ppg_export(self.h, _ptr(sizes), _ptr(t["snode"]), _ptr(t["dinfo"]), _ptr(t["quad"]),
                                            _ptr(t["rect"]), _ptr(t["sums"]), _ptr(t["counts"]), _ptr(t["bounds"])),
                    "cpt_ppg_export")
        return t

    def read_film(self, sums: bool = False, sumsq: bool = False):
        """rgb (h,w,3) float32; with sums/sumsq also the FP64 per-pixel sums and
        sums of squares of path radiance."""
        w, h = self._last
        rgb = np.zeros((h, w, 3), np.float32)
        s = np.zeros((h, w, 3), np.float64) if sums else None
        s2 = np.zeros((h, w, 3), np.float64) if sumsq else None
        self._check(self.lib.cpt_read_film(self.h, _ptr(rgb), _ptr(s), _ptr(s2)), "cpt_read_film")
        if sums and sumsq:
            return rgb, s, s2
        if sums:
            return rgb, s
        if sumsq:
            return rgb, s2
        return rgb


def _render_views(self, cameras, wh_spp, max_depth=32, seed=1, flags=0):
    """Render a batch of views in one streaming launch sequence.

    cameras: (n, 16) float32 camera records; wh_spp: (n, 3) int32 (width,
    height, spp_axis).  Returns (list of (h, w, 3) float32 images, per-view
    completion times in ms from the batch start, stats dict)."""
    cams = np.ascontiguousarray(cameras, np.float32).reshape(-1, 16)
    whs = np.ascontiguousarray(wh_spp, np.int32).reshape(-1, 3)
    n = cams.shape[0]
    px = int(np.sum(whs[:, 0].astype(np.int64) * whs[:, 1]))
    films = np.zeros(3 * px, np.float32)
    ms = np.zeros(n, np.float64)
    st = RenderStats()
    self._check(self.lib.cpt_render_views(self.h, n, _ptr(cams), _ptr(whs), max_depth, seed, flags,
                                          _ptr(films), _ptr(ms), C.byref(st)), "cpt_render_views")
    out, off = [], 0
    for w, h, _ in whs:
        out.append(films[off:off + 3 * w * h].reshape(h, w, 3))
        off += 3 * w * h
    return out, ms, st.as_dict()


Context.render_views = _render_views


def _guide_init(self, lr=1e-3, seed=1, guide_fraction=0.5, train_paths=0, train_steps=1, beta1=0.9, beta2=0.99,
                eps=1e-8, guide_depth=0, guide_target=0, infer_fp8=False, width=64, fraction_ramp=0):
    gp = GuideParams()
    gp.lr, gp.beta1, gp.beta2, gp.eps = lr, beta1, beta2, eps
    gp.seed, gp.guide_fraction, gp.train_paths, gp.train_steps = seed, guide_fraction, train_paths, train_steps
    gp.guide_depth = guide_depth
    gp.guide_target = guide_target
    gp.infer_fp8 = 1 if infer_fp8 else 0
    gp.width = width
    gp.fraction_ramp = fraction_ramp
    self._check(self.lib.cpt_guide_init(self.h, C.byref(gp)), "cpt_guide_init")


def _guide_shape(self):
    """(hidden width, lobes, outputs per vector, parameters) of the current guide."""
    sh = np.zeros(4, np.int32)
    self._check(self.lib.cpt_guide_shape(self.h, _ptr(sh)), "cpt_guide_shape")
    return tuple(int(v) for v in sh)


def _guide_params(self):
    out = np.zeros(self.guide_shape()[3], np.float32)
    self._check(self.lib.cpt_guide_get_params(self.h, _ptr(out)), "cpt_guide_get_params")
    return out


def _guide_set_params(self, params):
    p = np.ascontiguousarray(params, np.float32)
    if p.size != self.guide_shape()[3]:
        raise ValueError(f"guide has {self.guide_shape()[3]} parameters, got {p.size}")
    self._check(self.lib.cpt_guide_set_params(self.h, _ptr(p)), "cpt_guide_set_params")


def _guide_forward(self, inputs):
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 7)
    out = np.zeros((x.shape[0], self.guide_shape()[2]), np.float32)
    self._check(self.lib.cpt_guide_forward(self.h, _ptr(x), x.shape[0], _ptr(out)), "cpt_guide_forward")
    return out


def _unpack_mixtures(raw, lobes):
    h0 = raw[0].astype("<u4").view(np.float16).reshape(lobes, -1, 2).astype(np.float32)
    h1 = raw[1].astype("<u4").view(np.float16).reshape(lobes, -1, 2).astype(np.float32)
    return {"kmu_x": h0[..., 0], "kmu_y": h0[..., 1], "kmu_z": h1[..., 0], "weight": h1[..., 1],
            "c": raw[2].view(np.float32)}


def _debug_guide_mixtures(self, n):
    """Stored mixtures of the first n queue entries of the last guided bounce
    (cpt_debug_guide_mixtures), unpacked as guide_mixtures."""
    lobes = self.guide_shape()[1]
    raw = np.zeros((3, lobes, n), np.uint32)
    self._check(self.lib.cpt_debug_guide_mixtures(self.h, n, _ptr(raw)), "cpt_debug_guide_mixtures")
    return _unpack_mixtures(raw, lobes)


def _guide_mixtures(self, inputs):
    """Stored mixtures of the per-bounce inference code, unpacked: a dict of
    (lobes, n) arrays kmu_x, kmu_y, kmu_z, weight (the stored halves as float32)
    and c (float32); the pdf of w is sum_k exp(kmu_k . w + c_k)."""
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 7)
    lobes = self.guide_shape()[1]
    raw = np.zeros((3, lobes, x.shape[0]), np.uint32)
    self._check(self.lib.cpt_guide_mixtures(self.h, _ptr(x), x.shape[0], _ptr(raw)), "cpt_guide_mixtures")
    return _unpack_mixtures(raw, lobes)


def _guide_train_step(self, samples, apply=True, want_grad=False):
    s = np.ascontiguousarray(samples, np.float32).reshape(-1, 11)
    loss = np.zeros(1, np.float32)
    g = np.zeros(self.guide_shape()[3], np.float32) if want_grad else None
    self._check(self.lib.cpt_guide_train_step(self.h, _ptr(s), s.shape[0], 1 if apply else 0, _ptr(loss), _ptr(g)),
                "cpt_guide_train_step")
    return (float(loss[0]), g) if want_grad else float(loss[0])


Context.guide_init = _guide_init
Context.guide_shape = _guide_shape
Context.guide_params = _guide_params
Context.guide_set_params = _guide_set_params
Context.guide_forward = _guide_forward
Context.guide_mixtures = _guide_mixtures
Context.debug_guide_mixtures = _debug_guide_mixtures
Context.guide_train_step = _guide_train_step
