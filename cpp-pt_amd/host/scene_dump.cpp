// cpt_scene_dump: load a cpp-pt settings file with the host loader and write the
// flattened scene (the arrays cpt_upload_scene receives) as raw little-endian
// binary: header int32[5] = {nv, nt, nm, nl, spp}, then positions, normals,
// bitangents (nv*3 f32), indices (nt*3 i32), tri_material (nt i32), materials
// (nm*16 f32), light triangles (nl i32), camera (16 f32), env (3 f32), uvs (nv*2 f32),
// n_textures (i32), textures (n_textures*4 i32), n_texels (i32), texels (f32),
// material textures (nm*4 i32), n_point_lights (i32), point lights (n*6 f32).
// No device is used.
#include <cstdio>
#include <iostream>

#include "cpt_host.h"

using namespace cpppt;

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: cpt_scene_dump <settings> <out.bin>\n";
        return 2;
    }
    try {
        RenderSettings rs;
        Scene s;
        SceneData sd;
        Loader::load_render_settings(&rs, argv[1]);
        Loader::load_scene(&s, &sd, &rs, rs.scene_name);
        cpt_scene_desc d = s.desc();
        FILE* f = fopen(argv[2], "wb");
        if (!f) throw std::runtime_error("cannot write output");
        int32_t hdr[5] = {d.n_vertices, d.n_triangles, d.n_materials, d.n_lights, rs.spp};
        fwrite(hdr, sizeof(hdr), 1, f);
        fwrite(d.positions, sizeof(float), 3 * (size_t)d.n_vertices, f);
        fwrite(d.normals, sizeof(float), 3 * (size_t)d.n_vertices, f);
        fwrite(d.bitangents, sizeof(float), 3 * (size_t)d.n_vertices, f);
        fwrite(d.indices, sizeof(int32_t), 3 * (size_t)d.n_triangles, f);
        fwrite(d.tri_material, sizeof(int32_t), (size_t)d.n_triangles, f);
        fwrite(d.materials, sizeof(float), 16 * (size_t)d.n_materials, f);
        fwrite(d.light_triangles, sizeof(int32_t), (size_t)d.n_lights, f);
        fwrite(d.camera, sizeof(float), 16, f);
        fwrite(d.env, sizeof(float), 3, f);
        fwrite(s.uvs.data(), sizeof(float), s.uvs.size(), f);
        int32_t nt = d.n_textures, nx = (int32_t)s.texels.size();
        fwrite(&nt, sizeof(nt), 1, f);
        fwrite(s.textures.data(), sizeof(int32_t), s.textures.size(), f);
        fwrite(&nx, sizeof(nx), 1, f);
        fwrite(s.texels.data(), sizeof(float), s.texels.size(), f);
        fwrite(s.material_textures.data(), sizeof(int32_t), s.material_textures.size(), f);
        const int32_t np = (int32_t)(s.point_lights.size() / 6);
        fwrite(&np, sizeof(np), 1, f);
        fwrite(s.point_lights.data(), sizeof(float), s.point_lights.size(), f);
        fclose(f);
    } catch (const std::runtime_error& e) {
        std::cout << e.what() << std::endl;
        return 1;
    }
    return 0;
}
