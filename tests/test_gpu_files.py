"""The drop-in path through scene files: cpt_render (the reference's main,
src/main.cpp:89-146) loads .settings/.ptscene/.ptobj/.ptmat files with image
textures (loader.h:45-97) and a point light (loader.h:129-141), renders on the GPU and writes the 8-bit sRGB PNG
(rgb_image.h:300-318); the film equals an in-process render of the same
scene arrays."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import cpt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RENDER = os.path.join(ROOT, "cpp-pt_amd", "cpt_render")


def _read_png_rgb8(path):
    data = open(path, "rb").read()
    pos, idat, w = 8, b"", 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        t = data[pos + 4:pos + 8]
        if t == b"IHDR":
            w, h = struct.unpack(">II", data[pos + 8:pos + 16])
            assert data[pos + 16:pos + 18] == b"\x08\x02"
        elif t == b"IDAT":
            idat += data[pos + 8:pos + 8 + n]
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 3 * w + 1)
    assert (raw[:, 0] == 0).all()  # filter type 0 rows
    return raw[:, 1:].reshape(h, w, 3)


def _to_srgb8(x):
    """linear2srgb + put2char1 (rgb_image.h:23-35, 114-120) in float32/double."""
    x = np.asarray(x, np.float32)
    p = np.power(x, np.float32(1.0 / 2.4)).astype(np.float64)
    s = np.where(x > 0.0031308, ((1.0 + 0.055) * p - 0.055).astype(np.float32),
                 (12.92 * x.astype(np.float64)).astype(np.float32))
    s = np.clip(s, 0.0, np.float32(0.99999))
    return (s.astype(np.float64) * 255.0).astype(np.uint8)


@pytest.mark.parametrize("renderer", ["pathtracer_mis", "pure_pt"])
def test_cli_renders_textured_scene_files(ctx, tmp_path, renderer):
    W, H = 40, 32
    s = cpt.scenes.textured_room(res=(W, H))
    s.add_point_light((0.3, 0.4, 0.3), (1.0, 0.9, 0.7), 2.0)
    settings = s.write_cpppt(str(tmp_path), spp=2, renderer=renderer, max_depth=16, seed=7)
    print(f"cpt_render {renderer} {W}x{H}", flush=True)
    r = subprocess.run([RENDER, settings], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("s")
    png = _read_png_rgb8(str(tmp_path / "out.png"))
    assert png.shape == (H, W, 3)
    ctx.upload(s)
    ctx.render(W, H, 2, max_depth=16, seed=7, flags=cpt.CPT_RENDER_PURE_PT if renderer == "pure_pt" else 0)
    img = ctx.read_film()
    want = _to_srgb8(img)
    d = np.abs(png.astype(int) - want.astype(int))
    print(f"png vs in-process film: max |d| {d.max()}, equal {np.mean(d == 0) * 100:.2f}%")
    assert d.max() <= 1 and (d == 0).mean() > 0.99
    assert png.std() > 5  # not a blank image
