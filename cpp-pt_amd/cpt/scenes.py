"""Scene construction for the path tracer: flattened arrays + cpp-pt file I/O.

A Scene holds exactly what crosses the C ABI (include/cpt.h, cpt_scene_desc):
per-vertex positions/normals/bitangents, triangles in load order, a material
table (16 floats per material), emitter triangles in LightGroup order and a
camera frame.  `write_cpppt` writes the same scene in cpp-pt's own formats
(settings file, .ptscene, .ptobj, .ptmat: src/loader/loader.h:117-311), which the
C++ loader (cpp-pt_amd/host/loader.cpp) reads back to identical arrays.

Generators:
  box_room()      -- the 36-triangle closed box room with two rotated boxes and a
                     ceiling area light (512x512, FOV 39 deg)
  terrain_scene() -- 1024x512 fBm heightfield + 200 instanced icospheres
                     (subdivision 4), mixed Lambertian / GGX / dielectric
                     materials, 8 emissive quads and a constant environment
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

MAT_DIFFUSE, MAT_GLOSSY, MAT_EMISSIVE, MAT_DIELECTRIC, MAT_STANDARD = 0, 1, 2, 3, 4

f32 = np.float32


def material(kind, albedo=(0, 0, 0), rough=0.0, emission=(0, 0, 0), ior=1.5,
             normal_map=(0.5, 0.5, 1.0)):
    m = np.zeros(16, np.float32)
    m[0] = kind
    m[1:4] = albedo
    m[4] = rough
    m[5:8] = emission
    m[8] = ior
    m[9:12] = normal_map
    m[12] = 1.0 if kind in (MAT_DIFFUSE, MAT_GLOSSY) else 0.0
    return m


def _normalized(v):
    """cpp-pt normalized() on float32: v * float(1.0/double(float(sqrt(dot))))."""
    v = np.asarray(v, np.float32)
    d = f32(f32(v[0] * v[0]) + f32(v[1] * v[1]))
    d = f32(d + f32(v[2] * v[2]))
    ln = f32(math.sqrt(float(d)))
    inv = f32(1.0 / float(ln))
    return (v * inv).astype(np.float32)


def _cross(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array([f32(a[1] * b[2]) - f32(a[2] * b[1]),
                     f32(a[2] * b[0]) - f32(a[0] * b[2]),
                     f32(a[0] * b[1]) - f32(a[1] * b[0])], np.float32)


def camera_frame(origin, at, up, tan_fovy, res_x, res_y):
    """CameraPerspective set-up (camera_perspective.h:19-27, loader.h:117-128)."""
    o = np.asarray(origin, np.float32)
    fwd = _normalized(np.asarray(at, np.float32) - o)
    z = _normalized(fwd * f32(-1.0))
    x = _normalized(_cross(np.asarray(up, np.float32), z))
    y = _normalized(_cross(z, x))
    cam = np.zeros(16, np.float32)
    cam[0:3] = o
    cam[3:6] = x
    cam[6:9] = y
    cam[9:12] = z
    cam[12] = f32(tan_fovy)
    cam[13] = f32(f32(res_x) / f32(res_y))
    cam[14] = res_x
    cam[15] = res_y
    return cam


_LIBM = None


def _powf(x, y):
    """The C library's powf (what the C++ loader calls), for bit-identical sRGB decoding."""
    global _LIBM
    if _LIBM is None:
        import ctypes
        _LIBM = ctypes.CDLL("libm.so.6")
        _LIBM.powf.restype = ctypes.c_float
        _LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    return f32(_LIBM.powf(float(x), float(y)))


def srgb2linear(a):
    """rgb_image.h:123-129 on float32."""
    a = float(f32(a))
    if a > 0.04045:
        return _powf(f32((a + 0.055) / 1.055), f32(2.4))
    return f32(a / 12.92)


def image_texels(pixels, bw, srgb):
    """Image<T>(file, colour space) (rgb_image.h:161-197) from 8-bit samples (h, w, n):
    linear float32 texels, 3 channels (Vec3) or 1 (float)."""
    px = np.asarray(pixels, np.uint8)
    if px.ndim == 2:
        px = px[:, :, None]
    h, w, n = px.shape
    c = (px.astype(np.float64) / 255.0).astype(np.float32)  # char2float (rgb_image.h:19-21)
    if bw:
        out = c[:, :, :1]
    elif n >= 3:
        out = c[:, :, :3]
    else:
        out = np.repeat(c[:, :, :1], 3, axis=2)
    out = np.ascontiguousarray(out, np.float32)
    if srgb:
        flat = out.reshape(-1)
        lut = {}
        for i, v in enumerate(flat):
            k = float(v)
            if k not in lut:
                lut[k] = srgb2linear(k)
            flat[i] = lut[k]
    return out


@dataclass
class Mesh:
    positions: np.ndarray   # (nv,3) float32
    normals: np.ndarray     # (nv,3)
    tangents: np.ndarray    # (nv,3); bitangent = cross(normal, tangent) (loader.h:237)
    faces: np.ndarray       # (nf,3) int32
    material: int           # index into Scene.material_list
    uvs: np.ndarray = None  # (nv,2) texture coordinates (zeros when None)


@dataclass
class Scene:
    meshes: list = field(default_factory=list)
    material_list: list = field(default_factory=list)   # 16-float records
    camera_def: tuple = ((0, 0, 3.9), (0, 0, 0), (0, 1, 0), 0.35411, 512, 512)
    env: tuple = (0.0, 0.0, 0.0)
    images: list = field(default_factory=list)        # (pixels uint8 (h,w,n), bw, srgb)
    material_tex: list = field(default_factory=list)  # per material: albedo, roughness, metal, normal image
    point_lights: list = field(default_factory=list)  # (position, rgb colour, intensity)
    _flat: dict = None

    def add_material(self, rec, tex=(-1, -1, -1, -1)):
        """tex: image ids (add_image) for albedo, roughness, metal, normal map (-1: the
        record's constant).  An image metal makes the material MAT_STANDARD."""
        rec = np.asarray(rec, np.float32).copy()
        if tex[2] >= 0:
            rec[0] = MAT_STANDARD
        self.material_list.append(rec)
        self.material_tex.append(np.asarray(tex, np.int32))
        self._flat = None
        return len(self.material_list) - 1

    def add_image(self, pixels, bw, srgb):
        """An image texture from 8-bit samples (h, w) or (h, w, n); `bw` and `srgb` as the
        material slot reads it (albedo: Vec3, sRGB; roughness, metal: float, linear;
        normal map: Vec3, linear)."""
        self.images.append((np.ascontiguousarray(pixels, np.uint8), bool(bw), bool(srgb)))
        self._flat = None
        return len(self.images) - 1

    def add_mesh(self, positions, normals, tangents, faces, mat, uvs=None):
        self.meshes.append(Mesh(np.ascontiguousarray(positions, np.float32),
                                np.ascontiguousarray(normals, np.float32),
                                np.ascontiguousarray(tangents, np.float32),
                                np.ascontiguousarray(faces, np.int32), mat,
                                None if uvs is None else np.ascontiguousarray(uvs, np.float32)))
        self._flat = None

    def add_point_light(self, position, colour, intensity):
        """`light point x y z rgb(colour) intensity` (loader.h:129-141); intensity is
        the total power, radiance colour*intensity/(4 pi) (point_light.h:20-28)."""
        self.point_lights.append((tuple(float(x) for x in position), tuple(float(x) for x in colour), float(intensity)))
        self._flat = None

    def set_camera(self, origin, at, up, tan_fovy, res_x, res_y):
        self.camera_def = (tuple(origin), tuple(at), tuple(up), float(tan_fovy), int(res_x), int(res_y))
        self._flat = None

    @property
    def resolution(self):
        return self.camera_def[4], self.camera_def[5]

    def c_arrays(self):
        """Flattened arrays in load order (loader.h:353-367)."""
        if self._flat is not None:
            return self._flat
        pos, nrm, bit, idx, tmat, lights, uvs = [], [], [], [], [], [], []
        voff = 0
        toff = 0
        for m in self.meshes:
            pos.append(m.positions)
            nrm.append(m.normals)
            uvs.append(m.uvs if m.uvs is not None else np.zeros((len(m.positions), 2), np.float32))
            n, t = m.normals, m.tangents
            b = np.stack([n[:, 1] * t[:, 2] - n[:, 2] * t[:, 1],
                          n[:, 2] * t[:, 0] - n[:, 0] * t[:, 2],
                          n[:, 0] * t[:, 1] - n[:, 1] * t[:, 0]], axis=1).astype(np.float32)
            bit.append(b)
            idx.append(m.faces + voff)
            tmat.append(np.full(len(m.faces), m.material, np.int32))
            if self.material_list[m.material][0] == MAT_EMISSIVE:
                lights.append(np.arange(toff, toff + len(m.faces), dtype=np.int32))
            voff += len(m.positions)
            toff += len(m.faces)
        o, at, up, tf, rx, ry = self.camera_def
        self._flat = dict(
            positions=np.ascontiguousarray(np.concatenate(pos), np.float32),
            normals=np.ascontiguousarray(np.concatenate(nrm), np.float32),
            bitangents=np.ascontiguousarray(np.concatenate(bit), np.float32),
            indices=np.ascontiguousarray(np.concatenate(idx), np.int32),
            tri_material=np.ascontiguousarray(np.concatenate(tmat), np.int32),
            materials=np.ascontiguousarray(np.stack(self.material_list), np.float32),
            # LightGroup order: point lights (added while parsing), then emissive triangles
            light_tris=np.ascontiguousarray(np.concatenate(
                [-1 - np.arange(len(self.point_lights), dtype=np.int32)] + lights), np.int32),
            point_lights=self._point_light_table(),
            camera=camera_frame(o, at, up, tf, rx, ry),
            env=np.asarray(self.env, np.float32),
            uvs=np.ascontiguousarray(np.concatenate(uvs), np.float32),
        )
        texels, table = [], []
        off = 0
        for px, bw, srgb in self.images:
            t = image_texels(px, bw, srgb)
            table.append((t.shape[1], t.shape[0], t.shape[2], off))
            texels.append(t.reshape(-1))
            off += t.size
        mt = list(self.material_tex) + [np.full(4, -1, np.int32)] * (len(self.material_list) - len(self.material_tex))
        self._flat.update(
            textures=np.ascontiguousarray(np.array(table, np.int32).reshape(-1, 4)),
            texels=np.ascontiguousarray(np.concatenate(texels) if texels else np.zeros(1, np.float32), np.float32),
            material_textures=np.ascontiguousarray(np.stack(mt), np.int32),
        )
        return self._flat

    def _point_light_table(self):
        """Position and radiance per point light, radiance = (colour * intensity) *
        float(1.0 / float(4 pi)) in float32, as Vec3 / double evaluates (math.h:104-109)."""
        inv = f32(1.0 / float(f32(4.0 * math.pi)))
        rows = [list(p) + [float(f32(f32(c) * f32(i)) * inv) for c in col] for p, col, i in self.point_lights]
        return np.ascontiguousarray(np.array(rows, np.float32).reshape(-1, 6))

    @property
    def n_triangles(self):
        return sum(len(m.faces) for m in self.meshes)

    # ------------------------------------------------------------- cpp-pt files
    def write_cpppt(self, directory, name="scene", spp=1, output="out", renderer="pathtracer_mis",
                    max_depth=None, seed=None):
        """Write <dir>/<name>.settings, <name>.ptscene, mesh_*.ptobj, mat_*.ptmat."""
        os.makedirs(directory, exist_ok=True)
        o, at, up, tf, rx, ry = self.camera_def
        for k, (px, bw, srgb) in enumerate(self.images):
            write_png(os.path.join(directory, f"tex_{k}.png"), px)
        for k, rec in enumerate(self.material_list):
            tex = self.material_tex[k] if k < len(self.material_tex) else (-1, -1, -1, -1)

            def slot(i, const):
                return f"image(tex_{tex[i]}.png)" if tex[i] >= 0 else const
            with open(os.path.join(directory, f"mat_{k}.ptmat"), "w") as f:
                kind = int(rec[0])
                if kind == MAT_EMISSIVE:
                    f.write(f"EMISSION rgb({_n(rec[5])},{_n(rec[6])},{_n(rec[7])}) float(1.0)\n")
                elif kind == MAT_DIELECTRIC:
                    f.write(f"DIELECTRIC float({_n(rec[8])})\n")
                else:
                    metal = "1.0" if kind == MAT_GLOSSY else "0.0"
                    f.write(f"DISNEY {slot(0, f'rgb({_n(rec[1])},{_n(rec[2])},{_n(rec[3])})')} float(0.5) "
                            f"{slot(1, f'float({_n(rec[4])})')} {slot(2, f'float({metal})')} float(0.0) float(1.5) "
                            f"{slot(3, f'rgb({_n(rec[9])},{_n(rec[10])},{_n(rec[11])})')} float(1.0)\n")
        lines = [f"camera {_fmt(o)} {_fmt(at)} {_fmt(up)} {_n(tf)}"]
        for p, col, i in self.point_lights:
            lines.append(f"light point {_fmt(p)} rgb({_n(col[0])},{_n(col[1])},{_n(col[2])}) {_n(i)}")
        if any(self.env):
            lines.append(f"environment {_fmt(self.env)}")
        for i, m in enumerate(self.meshes):
            fn = f"mesh_{i}.ptobj"
            with open(os.path.join(directory, fn), "w") as f:
                f.write(f"{len(m.positions)} {len(m.faces)} TRIANGLE_MESH\n")
                uv = m.uvs if m.uvs is not None else np.zeros((len(m.positions), 2), np.float32)
                for p, q, n, t in zip(m.positions, uv, m.normals, m.tangents):
                    f.write(f"{_fmt(p)} {_fmt(q)} {_fmt(n)} {_fmt(t)}\n")
                for fc in m.faces:
                    f.write(f"{fc[0]} {fc[1]} {fc[2]}\n")
            lines.append(f"obj {fn} mat_{m.material}.ptmat")
        scene_path = os.path.join(directory, f"{name}.ptscene")
        with open(scene_path, "w") as f:
            f.write("\n".join(lines) + "\n")
        settings = os.path.join(directory, f"{name}.settings")
        with open(settings, "w") as f:
            f.write(f"renderer {renderer}\nspp {spp}\nresolution {rx} {ry}\nscene {scene_path}\n"
                    f"output {os.path.join(directory, output)}\n")
            if max_depth is not None:
                f.write(f"max_depth {max_depth}\n")
            if seed is not None:
                f.write(f"seed {seed}\n")
        return settings


def _n(x):
    return repr(float(np.float32(x)))


def write_png(path, pixels):
    """8-bit PNG (gray, gray+alpha, RGB or RGBA by the channel count), filter 0."""
    import struct
    import zlib
    px = np.asarray(pixels, np.uint8)
    if px.ndim == 2:
        px = px[:, :, None]
    h, w, n = px.shape
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[n]
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)) +
                chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def _fmt(v):
    return " ".join(_n(x) for x in v)


# --------------------------------------------------------------------- geometry helpers
def _quad(p0, p1, p2, p3, normal, tangent):
    """Quad p0..p3 (counter-clockwise seen from the normal side) as 2 triangles."""
    P = np.array([p0, p1, p2, p3], np.float32)
    N = np.tile(np.asarray(normal, np.float32), (4, 1))
    T = np.tile(np.asarray(tangent, np.float32), (4, 1))
    F = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return P, N, T, F


def _box(center, size, angle_deg):
    """Axis-aligned box rotated about y; outward normals; 6 faces, 12 triangles."""
    cx, cy, cz = center
    sx, sy, sz = (s * 0.5 for s in size)
    a = math.radians(angle_deg)
    ca, sa = math.cos(a), math.sin(a)

    def rot(p):
        x, y, z = p
        return (cx + ca * x + sa * z, cy + y, cz - sa * x + ca * z)

    def rotn(n):
        x, y, z = n
        return (ca * x + sa * z, y, -sa * x + ca * z)

    faces = [
        # (normal, tangent, corners ccw seen from outside)
        ((0, 1, 0), (1, 0, 0), [(-sx, sy, sz), (sx, sy, sz), (sx, sy, -sz), (-sx, sy, -sz)]),
        ((0, -1, 0), (1, 0, 0), [(-sx, -sy, -sz), (sx, -sy, -sz), (sx, -sy, sz), (-sx, -sy, sz)]),
        ((1, 0, 0), (0, 0, -1), [(sx, -sy, sz), (sx, -sy, -sz), (sx, sy, -sz), (sx, sy, sz)]),
        ((-1, 0, 0), (0, 0, 1), [(-sx, -sy, -sz), (-sx, -sy, sz), (-sx, sy, sz), (-sx, sy, -sz)]),
        ((0, 0, 1), (1, 0, 0), [(-sx, -sy, sz), (sx, -sy, sz), (sx, sy, sz), (-sx, sy, sz)]),
        ((0, 0, -1), (-1, 0, 0), [(sx, -sy, -sz), (-sx, -sy, -sz), (-sx, sy, -sz), (sx, sy, -sz)]),
    ]
    Ps, Ns, Ts, Fs = [], [], [], []
    for k, (n, t, cs) in enumerate(faces):
        P, N, T, F = _quad(*[rot(c) for c in cs], rotn(n), rotn(t))
        Ps.append(P)
        Ns.append(N)
        Ts.append(T)
        Fs.append(F + 4 * k)
    return np.concatenate(Ps), np.concatenate(Ns), np.concatenate(Ts), np.concatenate(Fs)


def box_room(res=(512, 512), light_gap=1e-3):
    """Closed 2x2x2 box room (open front), 5 Lambertian quads, two boxes rotated
    +/-17 deg (heights 0.6 and 1.2), one 0.47x0.38 ceiling light (17,12,4)
    `light_gap` below the ceiling."""
    s = Scene()
    white = s.add_material(material(MAT_DIFFUSE, albedo=(0.73, 0.73, 0.73)))
    red = s.add_material(material(MAT_DIFFUSE, albedo=(0.63, 0.07, 0.05)))
    green = s.add_material(material(MAT_DIFFUSE, albedo=(0.14, 0.45, 0.09)))
    light = s.add_material(material(MAT_EMISSIVE, emission=(17.0, 12.0, 4.0)))
    walls = [
        # floor y=-1 (normal +y)
        ((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1), (0, 1, 0), (1, 0, 0), white),
        # ceiling y=+1 (normal -y)
        ((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), (0, -1, 0), (1, 0, 0), white),
        # back wall z=-1 (normal +z)
        ((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (0, 0, 1), (1, 0, 0), white),
        # left wall x=-1 (normal +x)
        ((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (1, 0, 0), (0, 0, -1), red),
        # right wall x=+1 (normal -x)
        ((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1), (-1, 0, 0), (0, 0, 1), green),
    ]
    for p0, p1, p2, p3, n, t, m in walls:
        P, N, T, F = _quad(p0, p1, p2, p3, n, t)
        s.add_mesh(P, N, T, F, m)
    P, N, T, F = _box((0.33, -1.0 + 0.3, 0.35), (0.6, 0.6, 0.6), -17.0)
    s.add_mesh(P, N, T, F, white)
    P, N, T, F = _box((-0.33, -1.0 + 0.6, -0.3), (0.6, 1.2, 0.6), 17.0)
    s.add_mesh(P, N, T, F, white)
    y = 1.0 - light_gap
    hx, hz = 0.47 / 2, 0.38 / 2
    P, N, T, F = _quad((-hx, y, -hz), (hx, y, -hz), (hx, y, hz), (-hx, y, hz), (0, -1, 0), (1, 0, 0))
    s.add_mesh(P, N, T, F, light)
    s.set_camera((0.0, 0.0, 3.9), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.tan(math.radians(19.5)), *res)
    return s


def emissive_box(res=(64, 64), le=(2.0, 3.0, 5.0)):
    """Camera inside a closed box whose 6 faces all emit `le`: every camera ray
    hits an emitter first, so the image equals le exactly (emitters do not
    reflect, pathtracerMIS.h:67)."""
    s = Scene()
    em = s.add_material(material(MAT_EMISSIVE, emission=le))
    P, N, T, F = _box((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), 0.0)
    N = -N  # inward
    F = F[:, ::-1].copy()
    s.add_mesh(P, N, T, F, em)
    s.set_camera((0.1, 0.2, 0.3), (1.0, 0.5, -2.0), (0.0, 1.0, 0.0), 0.5, *res)
    return s


def single_quad_light(res=(32, 32), size=0.5, height=1.0, le=(1.0, 1.0, 1.0), albedo=0.5):
    """Diffuse floor (large) under a square emitter of side `size` at `height`,
    facing down; used for the one-bounce NEE closed form."""
    s = Scene()
    floor = s.add_material(material(MAT_DIFFUSE, albedo=(albedo, albedo, albedo)))
    em = s.add_material(material(MAT_EMISSIVE, emission=le))
    L = 50.0
    P, N, T, F = _quad((-L, 0, L), (L, 0, L), (L, 0, -L), (-L, 0, -L), (0, 1, 0), (1, 0, 0))
    s.add_mesh(P, N, T, F, floor)
    h = size / 2
    P, N, T, F = _quad((-h, height, -h), (h, height, -h), (h, height, h), (-h, height, h), (0, -1, 0), (1, 0, 0))
    s.add_mesh(P, N, T, F, em)
    s.set_camera((0.0, 0.5, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), 0.001, *res)
    return s


# --------------------------------------------------------------------- procedural terrain scene
def _value_noise(x, z, rng_table, period=256):
    """2-D value noise on an integer lattice with smoothstep interpolation."""
    xi = np.floor(x).astype(np.int64)
    zi = np.floor(z).astype(np.int64)
    fx = x - xi
    fz = z - zi
    ux = fx * fx * (3 - 2 * fx)
    uz = fz * fz * (3 - 2 * fz)

    def h(i, j):
        return rng_table[((i % period) * period + (j % period)) % (period * period)]

    a, b = h(xi, zi), h(xi + 1, zi)
    c, d = h(xi, zi + 1), h(xi + 1, zi + 1)
    return (a * (1 - ux) + b * ux) * (1 - uz) + (c * (1 - ux) + d * ux) * uz


def fbm_heightfield(nx=1024, nz=512, extent=(40.0, 20.0), octaves=6, amplitude=2.5, seed=2):
    """fBm heightfield (octaves of value noise, lacunarity 2, gain 0.5)."""
    rng = np.random.default_rng(seed)
    table = rng.random(256 * 256) * 2.0 - 1.0
    xs = np.linspace(-extent[0] / 2, extent[0] / 2, nx)
    zs = np.linspace(-extent[1] / 2, extent[1] / 2, nz)
    X, Z = np.meshgrid(xs, zs, indexing="xy")  # (nz, nx)
    Hh = np.zeros_like(X)
    freq, amp = 0.15, 1.0
    for o in range(octaves):
        off = rng.random(2) * 100.0
        Hh += amp * _value_noise(X * freq + off[0], Z * freq + off[1], table)
        freq *= 2.0
        amp *= 0.5
    return xs, zs, (Hh * amplitude).astype(np.float64)


def _heightfield_mesh(xs, zs, Hh):
    nz, nx = Hh.shape
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    P = np.stack([X, Hh, Z], axis=-1).reshape(-1, 3)
    dhdx = np.gradient(Hh, xs, axis=1)
    dhdz = np.gradient(Hh, zs, axis=0)
    N = np.stack([-dhdx, np.ones_like(Hh), -dhdz], axis=-1).reshape(-1, 3)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    T = np.stack([np.ones_like(Hh), dhdx, np.zeros_like(Hh)], axis=-1).reshape(-1, 3)
    T -= N * np.sum(T * N, axis=1, keepdims=True)
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    i = np.arange(nz - 1)[:, None] * nx + np.arange(nx - 1)[None, :]
    i = i.reshape(-1)
    # counter-clockwise seen from +y: (i, i+nx, i+1) and (i+1, i+nx, i+nx+1)
    F = np.concatenate([np.stack([i, i + nx, i + 1], 1), np.stack([i + 1, i + nx, i + nx + 1], 1)])
    F = F.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3)
    return P.astype(np.float32), N.astype(np.float32), T.astype(np.float32), F.astype(np.int32)


def icosphere(subdiv=4):
    """Unit icosphere: 20*4^subdiv faces (5120 at subdivision 4)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [tuple(np.asarray(v) / np.linalg.norm(v)) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (np.asarray(verts[a]) + np.asarray(verts[b])) * 0.5
                verts.append(tuple(m / np.linalg.norm(m)))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.asarray(verts, np.float64)
    F = np.asarray(faces, np.int32)
    # orient faces outward
    n = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    flip = np.sum(n * P[F].mean(axis=1), axis=1) < 0
    F[flip] = F[flip][:, ::-1]
    return P, F


def _sphere_tangents(N):
    up = np.where(np.abs(N[:, 1:2]) < 0.99, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    T = np.cross(up, N)
    return T / np.linalg.norm(T, axis=1, keepdims=True)


def terrain_scene(res=(1920, 1080), n_spheres=200, grid=(1024, 512), subdiv=4, seed=2, material_seed=3, n_lights=8):
    """Procedural ~2.07M-triangle scene: 1024x512 fBm heightfield (6 octaves),
    `n_spheres` instanced icospheres (subdivision 4) with radius U[0.05,0.5]
    resting on the terrain, materials 60% Lambertian / 25% GGX (roughness
    U[0.05,0.6]) / 15% dielectric (IOR 1.5), 8 emissive quads and a constant
    environment."""
    s = Scene()
    extent = (40.0, 20.0)
    xs, zs, Hh = fbm_heightfield(grid[0], grid[1], extent, seed=seed)
    ground = s.add_material(material(MAT_DIFFUSE, albedo=(0.55, 0.5, 0.42)))
    P, N, T, F = _heightfield_mesh(xs, zs, Hh)
    s.add_mesh(P, N, T, F, ground)
    rng = np.random.default_rng(material_seed)
    SP, SF = icosphere(subdiv)
    SN = SP.copy()
    ST = _sphere_tangents(SN)
    kinds = rng.random(n_spheres)
    for k in range(n_spheres):
        r = rng.uniform(0.05, 0.5)
        cx = rng.uniform(-extent[0] * 0.45, extent[0] * 0.45)
        cz = rng.uniform(-extent[1] * 0.45, extent[1] * 0.45)
        ix = int(np.clip(np.searchsorted(xs, cx), 0, len(xs) - 1))
        iz = int(np.clip(np.searchsorted(zs, cz), 0, len(zs) - 1))
        cy = Hh[iz, ix] + r * 0.8
        if kinds[k] < 0.60:
            m = s.add_material(material(MAT_DIFFUSE, albedo=tuple(rng.uniform(0.1, 0.9, 3))))
        elif kinds[k] < 0.85:
            m = s.add_material(material(MAT_GLOSSY, rough=float(rng.uniform(0.05, 0.6))))
        else:
            m = s.add_material(material(MAT_DIELECTRIC, ior=1.5))
        s.add_mesh(SP * r + np.array([cx, cy, cz]), SN, ST, SF, m)
    light = s.add_material(material(MAT_EMISSIVE, emission=(12.0, 11.0, 9.0)))
    for k in range(n_lights):
        lx = -extent[0] * 0.4 + k * (extent[0] * 0.8 / 7)
        lz = -2.0 + 4.0 * (k % 2)
        y = 9.0
        h = 0.75
        P, N, T, F = _quad((lx - h, y, lz - h), (lx + h, y, lz - h), (lx + h, y, lz + h), (lx - h, y, lz + h),
                           (0, -1, 0), (1, 0, 0))
        s.add_mesh(P, N, T, F, light)
    s.env = (0.25, 0.3, 0.4)
    s.set_camera((0.0, 7.0, 17.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.tan(math.radians(25.0)), *res)
    return s


# --------------------------------------------------------------------- hard-lighting interior
def _subdivided_boxes(centers, sizes, angles, sub=2):
    """Boxes resting on their centres' footprint, rotated about y; every face split
    into sub x sub quads (8 triangles per face for sub=2, 48 per box)."""
    g = np.linspace(0.0, 1.0, sub + 1)
    U, V = np.meshgrid(g, g, indexing="xy")
    U, V = U.ravel(), V.ravel()
    # unit cube faces: corner, u axis, v axis, outward normal (u x v = normal)
    faces = [((0, 1, 0), (1, 0, 0), (0, 0, 1), (0, 1, 0)),
             ((0, 0, 1), (1, 0, 0), (0, 0, -1), (0, -1, 0)),
             ((1, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0)),
             ((0, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0)),
             ((0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, 1)),
             ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, -1))]
    P, N, T = [], [], []
    for c0, u, v, n in faces:
        c0, u, v, n = (np.asarray(x, float) for x in (c0, u, v, n))
        if np.dot(np.cross(u, v), n) < 0:
            u, v = v, u
        P.append(c0[None] + U[:, None] * u[None] + V[:, None] * v[None])
        N.append(np.tile(n, (len(U), 1)))
        T.append(np.tile(u, (len(U), 1)))
    P = np.concatenate(P) - np.array([0.5, 0.0, 0.5])  # (54,3), base centred on the floor point
    N = np.concatenate(N)
    T = np.concatenate(T)
    nv = (sub + 1) ** 2
    quads = []
    for j in range(sub):
        for i in range(sub):
            a = j * (sub + 1) + i
            quads += [(a, a + 1, a + sub + 2), (a, a + sub + 2, a + sub + 1)]
    F1 = np.array(quads)
    F = np.concatenate([F1 + f * nv for f in range(6)])
    nb = len(centers)
    ca, sa = np.cos(angles), np.sin(angles)
    R = np.zeros((nb, 3, 3))
    R[:, 0, 0], R[:, 0, 2], R[:, 1, 1], R[:, 2, 0], R[:, 2, 2] = ca, sa, 1.0, -sa, ca
    Pw = np.einsum("bij,bvj->bvi", R, P[None] * sizes[:, None, :]) + centers[:, None, :]
    Nw = np.einsum("bij,vj->bvi", R, N)
    Tw = np.einsum("bij,vj->bvi", R, T)
    Fw = F[None] + (np.arange(nb) * len(P))[:, None, None]
    return Pw.reshape(-1, 3), Nw.reshape(-1, 3), Tw.reshape(-1, 3), Fw.reshape(-1, 3)


def interior_scene(res=(1280, 720), n_boxes=40000, seed=5, light_radiance=3000.0):
    """Hard lighting: a 10 x 3 x 10 room cluttered with `n_boxes` furniture boxes
    (48 triangles each), lit only by a 0.2 x 0.2 light in a corridor outside,
    through a 0.3-wide, 2-high door gap."""
    s = Scene()
    rng = np.random.default_rng(seed)
    wall = s.add_material(material(MAT_DIFFUSE, albedo=(0.7, 0.7, 0.7)))
    X0, X1, Y1, Z = 5.0, 8.0, 3.0, 5.0
    gap, gh = 0.15, 2.0
    quads = [
        # room: floor, ceiling, back wall x=-5, side walls z=+-5
        ((-X0, 0, Z), (X0, 0, Z), (X0, 0, -Z), (-X0, 0, -Z), (0, 1, 0)),
        ((-X0, Y1, -Z), (X0, Y1, -Z), (X0, Y1, Z), (-X0, Y1, Z), (0, -1, 0)),
        ((-X0, 0, Z), (-X0, 0, -Z), (-X0, Y1, -Z), (-X0, Y1, Z), (1, 0, 0)),
        ((X0, 0, -Z), (-X0, 0, -Z), (-X0, Y1, -Z), (X0, Y1, -Z), (0, 0, 1)),
        ((-X0, 0, Z), (X0, 0, Z), (X0, Y1, Z), (-X0, Y1, Z), (0, 0, -1)),
        # wall x=5 with the door gap (three pieces)
        ((X0, 0, -Z), (X0, 0, -gap), (X0, Y1, -gap), (X0, Y1, -Z), (-1, 0, 0)),
        ((X0, 0, gap), (X0, 0, Z), (X0, Y1, Z), (X0, Y1, gap), (-1, 0, 0)),
        ((X0, gh, -gap), (X0, gh, gap), (X0, Y1, gap), (X0, Y1, -gap), (-1, 0, 0)),
        # corridor outside: floor, ceiling, far wall x=8, ends z=+-5
        ((X0, 0, Z), (X1, 0, Z), (X1, 0, -Z), (X0, 0, -Z), (0, 1, 0)),
        ((X0, Y1, -Z), (X1, Y1, -Z), (X1, Y1, Z), (X0, Y1, Z), (0, -1, 0)),
        ((X1, 0, -Z), (X1, 0, Z), (X1, Y1, Z), (X1, Y1, -Z), (-1, 0, 0)),
        ((X1, 0, -Z), (X0, 0, -Z), (X0, Y1, -Z), (X1, Y1, -Z), (0, 0, 1)),
        ((X0, 0, Z), (X1, 0, Z), (X1, Y1, Z), (X0, Y1, Z), (0, 0, -1)),
    ]
    for p0, p1, p2, p3, n in quads:
        P = np.array([p0, p1, p2, p3], float)
        nn = np.asarray(n, float)
        g = np.cross(P[1] - P[0], P[2] - P[0])
        if np.dot(g, nn) < 0:
            P = P[[0, 3, 2, 1]]
        t = P[1] - P[0]
        t /= np.linalg.norm(t)
        s.add_mesh(P, np.tile(nn, (4, 1)), np.tile(t, (4, 1)), np.array([[0, 1, 2], [0, 2, 3]]), wall)
    # furniture
    cam = np.array([-4.0, 1.5, -3.0])
    cx = rng.uniform(-4.8, 4.6, 4 * n_boxes)
    cz = rng.uniform(-4.8, 4.8, 4 * n_boxes)
    keep = (np.hypot(cx - cam[0], cz - cam[2]) > 0.8) & ~((cx > 4.0) & (np.abs(cz) < 0.8))
    cx, cz = cx[keep][:n_boxes], cz[keep][:n_boxes]
    nb = len(cx)
    sizes = np.stack([rng.uniform(0.05, 0.25, nb), rng.uniform(0.05, 0.6, nb), rng.uniform(0.05, 0.25, nb)], 1)
    centers = np.stack([cx, np.zeros(nb), cz], 1)
    angles = rng.uniform(0, np.pi, nb)
    mats = [s.add_material(material(MAT_DIFFUSE, albedo=tuple(rng.uniform(0.2, 0.8, 3)))) for _ in range(6)]
    which = rng.integers(0, len(mats), nb)
    for k, m in enumerate(mats):
        sel = which == k
        if sel.any():
            P, N, T, F = _subdivided_boxes(centers[sel], sizes[sel], angles[sel])
            s.add_mesh(P, N, T, F, m)
    light = s.add_material(material(MAT_EMISSIVE, emission=(light_radiance,) * 3))
    h = 0.1
    P, N, T, F = _quad((X1 - 0.1, 1.5 - h, -h), (X1 - 0.1, 1.5 - h, h), (X1 - 0.1, 1.5 + h, h), (X1 - 0.1, 1.5 + h, -h),
                       (-1, 0, 0), (0, 0, 1))
    g = np.cross(P[1] - P[0], P[2] - P[0])
    if g[0] > 0:
        P = P[[0, 3, 2, 1]]
    s.add_mesh(P, N, T, F, light)
    s.set_camera(tuple(cam), (5.0, 1.0, 0.0), (0.0, 1.0, 0.0), math.tan(math.radians(30.0)), *res)
    return s


# --------------------------------------------------------------------- textured box room
def textured_room(res=(64, 64), tex_res=16, seed=9):
    """The box room with image textures: a checker albedo (sRGB) on the floor and
    back wall, a noise roughness map and a striped metal map on the tall box
    (diffuse where the stripes are dark, GGX where they are bright), a random
    tangent-space normal map on the left wall.  Texture coordinates span [0, 1.5]
    so that lookups wrap (REPEAT border)."""
    rng = np.random.default_rng(seed)
    s = box_room(res=res)
    n = tex_res
    yy, xx = np.mgrid[0:n, 0:n]
    checker = np.where(((xx // 4) + (yy // 4)) % 2 == 0, 230, 40).astype(np.uint8)
    albedo = np.stack([checker, (checker * 0.8).astype(np.uint8), (255 - checker // 2).astype(np.uint8)], axis=2)
    rough = rng.integers(20, 200, size=(n, n), dtype=np.uint8)
    metal = np.where((xx // 3) % 2 == 0, 0, 255).astype(np.uint8)
    nmap = np.stack([rng.integers(90, 166, (n, n)), rng.integers(90, 166, (n, n)),
                     np.full((n, n), 255)], axis=2).astype(np.uint8)
    i_alb = s.add_image(albedo, bw=False, srgb=True)
    i_rough = s.add_image(rough, bw=True, srgb=False)
    i_metal = s.add_image(metal, bw=True, srgb=False)
    i_nmap = s.add_image(nmap, bw=False, srgb=False)
    m_floor = s.add_material(material(MAT_DIFFUSE, albedo=(0.7, 0.7, 0.7)), tex=(i_alb, -1, -1, -1))
    m_box = s.add_material(material(MAT_DIFFUSE, albedo=(0.8, 0.75, 0.7), rough=0.3), tex=(-1, i_rough, i_metal, -1))
    m_wall = s.add_material(material(MAT_DIFFUSE, albedo=(0.6, 0.2, 0.2)), tex=(-1, -1, -1, i_nmap))

    def uv_of(P, axes):
        return np.stack([(P[:, axes[0]] + 1.0) * 0.75, (P[:, axes[1]] + 1.0) * 0.75], axis=1).astype(np.float32)
    # re-material the floor (mesh 0), back wall (2), left wall (3) and the tall box (6)
    for mi, mat, axes in ((0, m_floor, (0, 2)), (2, m_floor, (0, 1)), (3, m_wall, (2, 1)), (6, m_box, (0, 1))):
        m = s.meshes[mi]
        m.material = mat
        m.uvs = uv_of(m.positions, axes)
    s._flat = None
    return s


# --------------------------------------------------------------------- grid of interiors
def interior_grid_scene(res=(3840, 2160), grid=4, n_boxes=10000, seed=40, glossy_frac=0.4, lights_per_room=2,
                        light_radiance=3000.0):
    """Large-scale guided workload: grid x grid copies of the hard-lighting interior
    (interior_scene) with their own random furniture, a share `glossy_frac` of each
    room's furniture materials GGX with roughness U[0.02, 0.8], and `lights_per_room`
    small 0.2 x 0.2 emitters in each room's corridor (32 for the 4 x 4 grid).  The
    camera stands in the first room; rooms are 14 apart along x and 11 along z."""
    rng = np.random.default_rng(seed)
    out = Scene()
    for gi in range(grid):
        for gj in range(grid):
            room = interior_scene(res=res, n_boxes=n_boxes, seed=seed + 1 + gi * grid + gj)
            off = np.array([14.0 * gi, 0.0, 11.0 * gj], np.float32)
            base = len(out.material_list)
            furn = [k for k, m in enumerate(room.material_list) if m[0] == MAT_DIFFUSE][1:]  # not the walls
            for k, m in enumerate(room.material_list):
                rec = m.copy()
                if k in furn and rng.random() < glossy_frac:
                    rec[0] = MAT_GLOSSY
                    rec[4] = rng.uniform(0.02, 0.8)
                    rec[12] = 1.0
                out.add_material(rec)
            for m in room.meshes:
                if room.material_list[m.material][0] == MAT_EMISSIVE:
                    continue
                out.add_mesh(m.positions + off, m.normals, m.tangents, m.faces, base + m.material)
            light = out.add_material(material(MAT_EMISSIVE, emission=(light_radiance,) * 3))
            h = 0.1
            for q in range(lights_per_room):
                z = -3.0 + 6.0 * (q + 0.5) / lights_per_room
                P, N, T, F = _quad((7.9, 1.5 - h, z - h), (7.9, 1.5 - h, z + h), (7.9, 1.5 + h, z + h),
                                   (7.9, 1.5 + h, z - h), (-1, 0, 0), (0, 0, 1))
                if np.cross(P[1] - P[0], P[2] - P[0])[0] > 0:
                    P = P[[0, 3, 2, 1]]
                out.add_mesh(P + off, N, T, F, light)
    out.set_camera((-4.0, 1.5, -3.0), (5.0, 1.0, 0.0), (0.0, 1.0, 0.0), math.tan(math.radians(30.0)), *res)
    return out
