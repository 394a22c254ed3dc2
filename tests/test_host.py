"""Host-side checks that need no GPU: the C ABI library, the cpp-pt file
loader, and the scene generators."""
import os
import re
import subprocess

import numpy as np
import pytest

import cpt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DUMP = os.path.join(ROOT, "cpp-pt_amd", "cpt_scene_dump")


def test_abi_exports_every_declared_function():
    header = open(os.path.join(ROOT, "include", "cpt.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(cpt_\w+)\s*\(", header, re.M))
    assert declared == set(cpt.ABI_FUNCTIONS)
    lib = cpt.load_library()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.cpt_version() == 1001


def _load_dump(path):
    raw = open(path, "rb").read()
    nv, nt, nm, nl, spp = np.frombuffer(raw[:20], np.int32)
    off = 20

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return a

    out = dict(positions=take(np.float32, 3 * nv).reshape(-1, 3), normals=take(np.float32, 3 * nv).reshape(-1, 3),
               bitangents=take(np.float32, 3 * nv).reshape(-1, 3), indices=take(np.int32, 3 * nt).reshape(-1, 3),
               tri_material=take(np.int32, nt), materials=take(np.float32, 16 * nm).reshape(-1, 16),
               light_tris=take(np.int32, nl), camera=take(np.float32, 16), env=take(np.float32, 3))
    out["uvs"] = take(np.float32, 2 * nv).reshape(-1, 2)
    n_tex = int(take(np.int32, 1)[0])
    out["textures"] = take(np.int32, 4 * n_tex).reshape(-1, 4)
    n_texels = int(take(np.int32, 1)[0])
    out["texels"] = take(np.float32, n_texels)
    out["material_textures"] = take(np.int32, 4 * nm).reshape(-1, 4)
    n_point = int(take(np.int32, 1)[0])
    out["point_lights"] = take(np.float32, 6 * n_point).reshape(-1, 6)
    assert off == len(raw)
    return out, int(spp)


@pytest.mark.parametrize("make", [lambda: cpt.scenes.box_room(res=(640, 480)),
                                  lambda: cpt.scenes.terrain_scene(n_spheres=12, grid=(20, 10), subdiv=1)])
def test_loader_reads_generated_files_bit_exactly(tmp_path, make):
    s = make()
    settings = s.write_cpppt(str(tmp_path), spp=3)
    out = tmp_path / "dump.bin"
    r = subprocess.run([DUMP, settings, str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got, spp = _load_dump(str(out))
    assert spp == 3
    want = s.c_arrays()
    for k in ("positions", "normals", "bitangents", "indices", "tri_material", "light_tris", "camera", "env"):
        assert np.array_equal(got[k], want[k]), k
    # materials: the loader deduplicates nothing here (one file per material)
    assert np.array_equal(got["materials"][:, :13], want["materials"][:, :13])


def test_loader_reads_image_textures(tmp_path):
    """Image textures through the file loader (loader.h:45-97, rgb_image.h:161-197):
    PNG decode, char2float, the sRGB decode of albedo images, the channel choice of
    float (first channel) and Vec3 images, texture coordinates, and per-material
    texture slots.  The loader numbers materials and images in order of first use,
    so the comparison goes through each triangle's material."""
    s = cpt.scenes.textured_room(res=(32, 32))
    settings = s.write_cpppt(str(tmp_path), spp=2)
    out = tmp_path / "dump.bin"
    r = subprocess.run([DUMP, settings, str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got, _ = _load_dump(str(out))
    want = s.c_arrays()
    for k in ("positions", "normals", "bitangents", "indices", "light_tris", "camera", "env", "uvs"):
        assert np.array_equal(got[k], want[k]), k
    assert len(got["textures"]) == len(s.images) == 4

    def image(a, m, slot):
        t = a["material_textures"][m, slot]
        if t < 0:
            return None
        w, h, ch, o = a["textures"][t]
        return (w, h, ch), a["texels"][o:o + w * h * ch]
    fields = {0: slice(1, 4), 1: slice(4, 5), 3: slice(9, 12)}
    pairs = set(zip(got["tri_material"].tolist(), want["tri_material"].tolist()))
    assert len(pairs) == len(set(want["tri_material"].tolist()))
    n_tex = 0
    for gm, wm in pairs:
        assert got["materials"][gm, 0] == want["materials"][wm, 0]
        for slot in range(4):
            gi, wi = image(got, gm, slot), image(want, wm, slot)
            assert (gi is None) == (wi is None), (wm, slot)
            if gi is None:
                if slot in fields:
                    f = fields[slot]
                    assert np.array_equal(got["materials"][gm, f], want["materials"][wm, f]), (wm, slot)
                continue
            n_tex += 1
            assert gi[0] == wi[0] and np.array_equal(gi[1], wi[1]), (wm, slot)
    assert n_tex == 4
    # sRGB albedo: 230/255 decodes to the IEC 61966-2-1 value
    alb = image(want, 4, 0)[1]
    c = 230 / 255
    assert abs(alb[0] - ((c + 0.055) / 1.055) ** 2.4) < 1e-6


def _write_blender_layout(s, d):
    """The layout of the reference's Blender exporter (scripts/blender_export_cpppt.py):
    models/, materials/ and images/ below the scene directory, one material token
    per line, floats printed by Python's str(), point lights as `light point`."""
    for sub in ("models", "materials", "images"):
        os.makedirs(os.path.join(d, sub), exist_ok=True)
    for k, (px, bw, srgb) in enumerate(s.images):
        cpt.scenes.write_png(os.path.join(d, "images", f"img{k}.png"), px)

    def f(x):
        return str(float(x))
    lines = []
    for k, rec in enumerate(s.material_list):
        tex = s.material_tex[k]
        kind = int(rec[0])
        with open(os.path.join(d, "materials", f"M{k}.ptmat"), "w") as fm:
            if kind == cpt.scenes.MAT_EMISSIVE:
                fm.write(f"EMISSION\nrgb({f(rec[5])},{f(rec[6])},{f(rec[7])})\nfloat(1.0)\n")
                continue
            slot = [f"image(images/img{t}.png)" if t >= 0 else None for t in tex]
            metal = "1.0" if kind == cpt.scenes.MAT_GLOSSY else "0.0"
            toks = [slot[0] or f"rgb({f(rec[1])},{f(rec[2])},{f(rec[3])})", "float(0.5)",
                    slot[1] or f"float({f(rec[4])})", slot[2] or f"float({metal})", "float(0.0)", "float(1.3)",
                    slot[3] or "rgb(0.5,0.5,1.0)", "float(1.0)"]
            fm.write("DISNEY\n" + "\n".join(toks) + "\n")
    for i, m in enumerate(s.meshes):
        with open(os.path.join(d, "models", f"Obj{i}.ptobj"), "w") as fo:
            fo.write(f"{len(m.positions)} {len(m.faces)} TRIANGLE_MESH\n")
            uv = m.uvs if m.uvs is not None else np.zeros((len(m.positions), 2), np.float32)
            for p, q, n, t in zip(m.positions, uv, m.normals, m.tangents):
                fo.write(" ".join(f(x) for x in (*p, *q, *n, *t)) + "\n")
            for fc in m.faces:
                fo.write(f"{fc[0]} {fc[1]} {fc[2]}\n")
        lines.append(f"obj models/Obj{i}.ptobj materials/M{m.material}.ptmat")
    for p, col, power in s.point_lights:
        lines.append(f"light point {f(p[0])} {f(p[1])} {f(p[2])} rgb({f(col[0])},{f(col[1])},{f(col[2])}) {f(power)}")
    o, at, up, tf, rx, ry = s.camera_def
    lines.append("camera " + " ".join(f(x) for x in (*o, *at, *up, tf)))
    with open(os.path.join(d, "scene.ptscene"), "w") as fs:
        fs.write("\n".join(lines) + "\n")
    with open(os.path.join(d, "scene.settings"), "w") as fs:
        fs.write(f"renderer pathtracer_mis\nspp 1\nresolution {rx} {ry}\nscene {os.path.join(d, 'scene.ptscene')}\n"
                 f"output {os.path.join(d, 'out')}\n")
    return os.path.join(d, "scene.settings")


def test_loader_reads_blender_export_layout_and_point_lights(tmp_path):
    """A textured room with two point lights written the way the reference's Blender
    exporter writes scenes: geometry, uvs, textures and materials load as from
    the generator, and the LightGroup holds the point lights first (they are added
    while the .ptscene is parsed, loader.h:129-141), then the emissive triangles."""
    s = cpt.scenes.textured_room(res=(24, 16))
    s.add_point_light((0.4, 0.6, 0.1), (1.0, 0.8, 0.6), 3.0)
    s.add_point_light((-0.5, 0.2, 0.5), (0.2, 0.4, 1.0), 1.5)
    settings = _write_blender_layout(s, str(tmp_path))
    r = subprocess.run([DUMP, settings, str(tmp_path / "d.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got, _ = _load_dump(str(tmp_path / "d.bin"))
    want = s.c_arrays()
    for k in ("positions", "normals", "bitangents", "indices", "uvs", "light_tris", "point_lights"):
        assert np.array_equal(got[k], want[k]), k
    assert got["light_tris"][:2].tolist() == [-1, -2]
    # radiance = colour * power / (4 pi) in float32
    assert abs(got["point_lights"][0, 3] - 3.0 / (4 * np.pi)) < 1e-6
    gm, wm = got["tri_material"], want["tri_material"]
    assert np.array_equal(got["materials"][gm, 0], want["materials"][wm, 0])
    assert len(got["textures"]) == 4
    # camera: the generator's frame
    assert np.allclose(got["camera"], want["camera"], atol=1e-6)


def test_point_light_colour_textures(tmp_path):
    """get_texture<false, true> for the point light's colour (loader.h:129-141):
    rgb(...) is taken as written, float(x) is decoded from sRGB, and an image is
    read at uv 0, which lands on its first texel."""
    s = cpt.scenes.box_room(res=(8, 8))
    settings = s.write_cpppt(str(tmp_path))
    cpt.scenes.write_png(str(tmp_path / "c.png"), np.array([[[200, 100, 50], [0, 0, 0]]], np.uint8))
    with open(tmp_path / "scene.ptscene", "a") as fs:
        fs.write("light point 0 0.5 0 float(0.5) 2.0\n")
        fs.write(f"light point 0 0.5 0 image({tmp_path / 'c.png'}) 2.0\n")
    r = subprocess.run([DUMP, settings, str(tmp_path / "d.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    got, _ = _load_dump(str(tmp_path / "d.bin"))
    inv = np.float32(1.0 / float(np.float32(4 * np.pi)))
    srgb_half = cpt.scenes.srgb2linear(np.float32(0.5))
    assert np.array_equal(got["point_lights"][0, 3:], np.full(3, np.float32(np.float32(srgb_half) * np.float32(2.0)) * inv))
    first = cpt.scenes.image_texels(np.array([[[200, 100, 50]]], np.uint8), False, True)[0, 0]
    assert np.array_equal(got["point_lights"][1, 3:], (first * np.float32(2.0)).astype(np.float32) * inv)
    bad = tmp_path / "bad"
    bad.mkdir()
    settings = s.write_cpppt(str(bad))
    with open(bad / "scene.ptscene", "a") as fs:
        fs.write("light spot 0 0 0 rgb(1,1,1) 1.0\n")
    r = subprocess.run([DUMP, settings, str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown light type: spot" in r.stdout


def test_png_reader_formats(tmp_path):
    """Gray, gray+alpha, RGB, RGBA, palette (with tRNS), 16-bit and 1/2/4-bit
    images, every filter type: the loader's texels equal the decoded samples."""
    rng = np.random.default_rng(4)
    cases = []
    px = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    cases.append(("rgb8", _png(px, 2, 8, filt=[0, 1, 2, 3, 4]), px))
    px = rng.integers(0, 256, (4, 6, 4), dtype=np.uint8)
    cases.append(("rgba8", _png(px, 6, 8, filt=[4, 3, 2, 1]), px[:, :, :3]))
    g = rng.integers(0, 256, (6, 5), dtype=np.uint8)
    cases.append(("gray8", _png(g, 0, 8, filt=[1, 4, 0, 3, 2, 4]), np.repeat(g[:, :, None], 3, axis=2)))
    ga = rng.integers(0, 256, (3, 9, 2), dtype=np.uint8)
    cases.append(("graya8", _png(ga, 4, 8, filt=[3]), np.repeat(ga[:, :, :1], 3, axis=2)))
    w16 = rng.integers(0, 65536, (4, 4, 3), dtype=np.uint16)
    cases.append(("rgb16", _png(w16, 2, 16, filt=[2, 4, 1, 3]), (w16 >> 8).astype(np.uint8)))
    pal = rng.integers(0, 256, (5, 3), dtype=np.uint8)
    idx = rng.integers(0, 5, (3, 11), dtype=np.uint8)
    cases.append(("pal8", _png(idx, 3, 8, filt=[1], plte=pal), pal[idx]))
    cases.append(("pal8trns", _png(idx, 3, 8, filt=[0], plte=pal, trns=bytes([7, 9])), pal[idx]))
    for depth in (1, 2, 4):
        v = rng.integers(0, 1 << depth, (3, 13), dtype=np.uint8)
        scale = 255 // ((1 << depth) - 1)
        cases.append((f"gray{depth}", _png(v, 0, depth, filt=[4, 0, 2]), np.repeat((v * scale)[:, :, None], 3, axis=2)))
        idx4 = rng.integers(0, min(5, 1 << depth), (2, 9), dtype=np.uint8)
        cases.append((f"pal{depth}", _png(idx4, 3, depth, filt=[0, 1], plte=pal), pal[idx4]))
    for name, data, want in cases:
        d = tmp_path / name
        d.mkdir()
        (d / "t.png").write_bytes(data)
        (d / "m.ptmat").write_text("DISNEY image(t.png) float(0.5) float(0.5) float(0.0) float(0.0) float(1.5) "
                                   "rgb(0.5,0.5,1.0) float(1.0)\n")
        (d / "q.ptobj").write_text("3 1 TRIANGLE_MESH\n0 0 0 0 0 0 0 1 1 0 0\n1 0 0 1 0 0 0 1 1 0 0\n"
                                   "0 1 0 0 1 0 0 1 1 0 0\n0 1 2\n")
        (d / "s.ptscene").write_text("camera 0 0 1 0 0 0 0 1 0 0.5\nobj q.ptobj m.ptmat\n")
        (d / "s.settings").write_text(f"resolution 4 4\nscene {d / 's.ptscene'}\n")
        r = subprocess.run([DUMP, str(d / "s.settings"), str(d / "x.bin")], capture_output=True, text=True)
        assert r.returncode == 0, name + r.stdout + r.stderr
        got, _ = _load_dump(str(d / "x.bin"))
        h, w = want.shape[:2]
        assert got["textures"].tolist() == [[w, h, 3, 0]], name
        assert np.array_equal(got["texels"], cpt.scenes.image_texels(want, False, True).reshape(-1)), name


def _png(px, ctype, depth, filt, plte=None, trns=None):
    """A PNG with the given colour type and bit depth; row y uses filter filt[y % len]."""
    import struct
    import zlib
    a = np.asarray(px)
    if a.ndim == 2:
        a = a[:, :, None]
    h, w, ch = a.shape
    rows = []
    for y in range(h):
        if depth == 16:
            row = a[y].astype(">u2").tobytes()
            bpp = 2 * ch
        elif depth == 8:
            row = a[y].astype(np.uint8).tobytes()
            bpp = ch
        else:
            per = 8 // depth
            b = bytearray((w + per - 1) // per)
            for x in range(w):
                b[x // per] |= int(a[y, x, 0]) << ((per - 1 - x % per) * depth)
            row = bytes(b)
            bpp = 1
        rows.append(row)
    raw = bytearray()
    prev = bytes(len(rows[0]))
    for y, row in enumerate(rows):
        ft = filt[y % len(filt)]
        out = bytearray(len(row))
        for i in range(len(row)):
            left = row[i - bpp] if i >= bpp else 0
            up = prev[i]
            ul = prev[i - bpp] if i >= bpp else 0
            if ft == 0:
                pred = 0
            elif ft == 1:
                pred = left
            elif ft == 2:
                pred = up
            elif ft == 3:
                pred = (left + up) >> 1
            else:
                p = left + up - ul
                pa, pb, pc = abs(p - left), abs(p - up), abs(p - ul)
                pred = left if pa <= pb and pa <= pc else (up if pb <= pc else ul)
            out[i] = (row[i] - pred) & 255
        raw += bytes([ft]) + out
        prev = row

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if plte is not None:
        data += chunk(b"PLTE", np.asarray(plte, np.uint8).tobytes())
    if trns is not None:
        data += chunk(b"tRNS", trns)
    # split the image data over two IDAT chunks
    z = zlib.compress(bytes(raw))
    return data + chunk(b"IDAT", z[:len(z) // 2]) + chunk(b"IDAT", z[len(z) // 2:]) + chunk(b"IEND", b"")


def test_loader_error_messages(tmp_path):
    bad = tmp_path / "bad.settings"
    bad.write_text("renderer pathtracer_mis\nfoo 3\n")
    r = subprocess.run([DUMP, str(bad), str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown render setting: foo" in r.stdout
    s = cpt.scenes.box_room()
    settings = s.write_cpppt(str(tmp_path / "sc"))
    # a missing image texture (loader.h:45-97 via rgb_image.h:161-197)
    (tmp_path / "sc" / "mat_0.ptmat").write_text(
        "DISNEY image(nope.png) float(0.5) float(0.5) float(0.0) float(0.0) float(1.5) rgb(0.5,0.5,1.0) float(1.0)\n")
    with open(os.path.join(str(tmp_path / "sc"), "scene.ptscene"), "w") as f:
        f.write("camera 0 0 3.9 0 0 0 0 1 0 0.35\nobj mesh_0.ptobj mat_0.ptmat\n")
    r = subprocess.run([DUMP, settings, str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "Cannot open image file" in r.stdout
    (tmp_path / "m.ptmat").write_text("PHONG rgb(1,1,1)\n")
    (tmp_path / "m.ptscene").write_text("camera 0 0 1 0 0 0 0 1 0 0.5\nobj missing.ptobj m.ptmat\n")
    (tmp_path / "m.settings").write_text(f"scene {tmp_path / 'm.ptscene'}\n")
    r = subprocess.run([DUMP, str(tmp_path / "m.settings"), str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "Insuficient number of lines" in r.stdout


def test_material_file_repeats_share_one_material(tmp_path):
    # parse_mat (loader.h:147-150): a repeated .ptmat name re-uses the material
    s = cpt.scenes.box_room()
    settings = s.write_cpppt(str(tmp_path))
    r = subprocess.run([DUMP, settings, str(tmp_path / "d.bin")], capture_output=True, text=True)
    assert r.returncode == 0
    got, _ = _load_dump(str(tmp_path / "d.bin"))
    assert got["materials"].shape[0] == 4  # white, red, green, light
    assert (got["tri_material"] == 0).sum() == 6 + 24  # floor, ceiling, back wall, 2 boxes


def test_box_room_shape():
    a = cpt.scenes.box_room().c_arrays()
    assert a["indices"].shape == (36, 3)
    assert a["light_tris"].tolist() == [34, 35]
    # inward-facing walls: shading normal agrees with the winding
    P, I, N = a["positions"], a["indices"], a["normals"]
    g = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    assert (np.sum(g * N[I[:, 0]], axis=1) > 0).all()


def test_terrain_scene_size():
    s = cpt.scenes.terrain_scene()
    a = s.c_arrays()
    assert a["indices"].shape[0] == 1023 * 511 * 2 + 200 * 5120 + 16
    kinds = a["materials"][:, 0].astype(int)
    assert set(kinds) == {0, 1, 2, 3}
    P, I, N = a["positions"], a["indices"], a["normals"]
    g = np.cross(P[I[:, 1]] - P[I[:, 0]], P[I[:, 2]] - P[I[:, 0]])
    assert (np.sum(g * N[I[:, 0]], axis=1) > 0).mean() > 0.999


@pytest.mark.parametrize("make", [lambda: cpt.scenes.box_room(res=(64, 48)), lambda: cpt.scenes.textured_room(res=(32, 32))])
def test_scene_file_load_c_abi(tmp_path, make):
    """cpt_scene_file_load: the C-ABI loader a reference-side renderer calls with
    its settings' scene_name gives the same flattened scene as the host loader's
    dump, field for field, with the camera set up at the resolution passed in."""
    s = make()
    settings = s.write_cpppt(str(tmp_path), spp=2)
    r = subprocess.run([DUMP, settings, str(tmp_path / "d.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    want, _ = _load_dump(str(tmp_path / "d.bin"))
    rx, ry = s.resolution
    got = cpt.load_scene_file(str(tmp_path / "scene.ptscene"), rx, ry)
    for k in ("positions", "normals", "bitangents", "indices", "tri_material", "materials", "light_tris",
              "camera", "env", "uvs", "textures", "texels", "material_textures", "point_lights"):
        assert np.array_equal(got[k].reshape(-1), want[k].reshape(-1)), k
    assert np.array_equal(got["positions"], s.c_arrays()["positions"])
    # another film size changes only the camera's aspect and resolution entries
    other = cpt.load_scene_file(str(tmp_path / "scene.ptscene"), 2 * rx, ry)
    assert np.array_equal(other["camera"][:13], got["camera"][:13])
    assert other["camera"][13] == np.float32(2 * rx) / np.float32(ry) and other["camera"][14] == 2 * rx


def test_scene_file_load_errors(tmp_path):
    with pytest.raises(cpt.CptError, match="Cannot open scene file"):
        cpt.load_scene_file(str(tmp_path / "missing.ptscene"), 8, 8)
    s = cpt.scenes.box_room(res=(8, 8))
    s.write_cpppt(str(tmp_path))
    with open(tmp_path / "scene.ptscene", "a") as fs:
        fs.write("sphere 0 0 0 1\n")
    with pytest.raises(cpt.CptError, match="Unknown scene obj: sphere"):
        cpt.load_scene_file(str(tmp_path / "scene.ptscene"), 8, 8)
    with pytest.raises(cpt.CptError, match="non-positive resolution"):
        cpt.load_scene_file(str(tmp_path / "scene.ptscene"), 0, 8)
