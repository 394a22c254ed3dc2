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
    assert lib.cpt_version() == 1000


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


def test_loader_error_messages(tmp_path):
    bad = tmp_path / "bad.settings"
    bad.write_text("renderer pathtracer_mis\nfoo 3\n")
    r = subprocess.run([DUMP, str(bad), str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown render setting: foo" in r.stdout
    s = cpt.scenes.box_room()
    settings = s.write_cpppt(str(tmp_path / "sc"))
    with open(os.path.join(str(tmp_path / "sc"), "scene.ptscene"), "a") as f:
        f.write("light point 0 0 0 rgb(1,1,1) 1.0\n")
    r = subprocess.run([DUMP, settings, str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "Point lights" in r.stdout
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
