// Settings / scene loading, camera set-up, flattening and image output for the
// host side of libcpt.  File formats and parsing rules follow
// src/loader/loader.h (settings: 256-311, scene: 314-374, .ptobj: 209-251,
// .ptmat: 146-207, textures: 29-97) and the image output of
// src/image/rgb_image.h:114-129,300-318.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <unordered_map>

#include "cpt_host.h"

namespace cpppt {

namespace {

Vec3f v(float x, float y, float z) {
    Vec3f r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
Vec3f cross(const Vec3f& a, const Vec3f& b) {
    return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// math.h:137-140 with operator/(T) = *(1.0/a) and sqrt in double
Vec3f normalized(const Vec3f& a) {
    float len = (float)std::sqrt((double)(a.x * a.x + a.y * a.y + a.z * a.z));
    float inv = (float)(1.0 / (double)len);
    return v(a.x * inv, a.y * inv, a.z * inv);
}

// rgb_image.h:122-129
float srgb2linear(float a) {
    if (a > 0.04045) return powf((float)(((double)a + 0.055) / 1.055), 2.4f);
    return (float)((double)a / 12.92);
}
// rgb_image.h:114-120
float linear2srgb(float a) {
    if (a > 0.0031308) return (float)((1.0 + 0.055) * (double)powf(a, (float)(1.0 / 2.4)) - 0.055);
    return (float)(12.92 * (double)a);
}
// rgb_image.h:23-35
unsigned char put2char1(float a) {
    if (a <= 0.0) a = 0.0f;
    if (a >= 0.99999) a = 0.99999f;
    return (unsigned char)(unsigned int)((double)a * 255.0);
}

// loader.h:29-31: "float(x)"
float parse_float_tex(const std::string& s) { return (float)atof(s.substr(6, s.length() - 7).c_str()); }

// loader.h:34-43: "rgb(r,g,b)"
void parse_rgb(const std::string& s, float* r, float* g, float* b) {
    size_t c1 = s.find(",");
    std::string s2 = s.substr(4, c1 - 4);
    *r = (float)atof(s2.c_str());
    s2 = s.substr(c1 + 1);
    size_t c2 = s2.find(",");
    *g = (float)atof(s2.substr(0, c2).c_str());
    *b = (float)atof(s2.substr(c2 + 1, s2.length() - c2 - 2).c_str());
}

struct Tex {
    Vec3f value;     // constant textures
    int image = -1;  // image textures: index into SceneData::images
};

// Image<T>(filename, colour space) (rgb_image.h:161-197): 8-bit samples / 255;
// Vec3 images take the first three channels (or the first one three times
// when there are fewer), float images the first; sRGB images are decoded
int load_image(SceneData* sd, const std::string& path, bool bw, bool srgb) {
    const std::string key = path + (bw ? "|bw" : "|rgb") + (srgb ? "|srgb" : "|linear");
    auto it = sd->image_index.find(key);
    if (it != sd->image_index.end()) return it->second;
    int w, h, n;
    std::vector<uint8_t> px;
    load_png(path, &w, &h, &n, &px);
    TexImage im;
    im.w = w;
    im.h = h;
    im.ch = bw ? 1 : 3;
    im.texels.resize((size_t)w * h * im.ch);
    auto c2f = [](uint8_t a) { return (float)((double)(float)a / 255.0); };  // char2float, rgb_image.h:19-21
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const uint8_t* s = &px[i * n];
        if (bw) {
            im.texels[i] = c2f(s[0]);
        } else {
            const bool rgb = n >= 3;
            im.texels[3 * i] = c2f(s[0]);
            im.texels[3 * i + 1] = c2f(rgb ? s[1] : s[0]);
            im.texels[3 * i + 2] = c2f(rgb ? s[2] : s[0]);
        }
    }
    if (srgb)
        for (float& t : im.texels) t = srgb2linear(t);
    sd->images.push_back(std::move(im));
    sd->image_index[key] = (int)sd->images.size() - 1;
    return (int)sd->images.size() - 1;
}

// loader.h:45-97; bw textures keep the first channel
Tex get_texture(SceneData* sd, const std::string& dir, const std::string& tex, bool bw, bool srgb) {
    Tex t;
    if (tex.substr(0, 3) == "rgb") {
        float r, g, b;
        parse_rgb(tex, &r, &g, &b);
        t.value = bw ? v(r, r, r) : v(r, g, b);  // rgb(...) is never sRGB-decoded (loader.h:51-55)
        return t;
    } else if (tex.substr(0, 5) == "float") {
        float r = parse_float_tex(tex);
        if (srgb) r = srgb2linear(r);
        t.value = v(r, r, r);
        return t;
    } else if (tex.substr(0, 5) == "image") {
        t.image = load_image(sd, dir + tex.substr(6, tex.length() - 7), bw, srgb);
        t.value = v(0, 0, 0);
        return t;
    }
    throw std::runtime_error("Unknown type of tex: " + tex);
}

std::string trim(const std::string& s, const std::string& ws = " \n\t\r") {
    size_t i = s.find_first_not_of(ws);
    if (i == std::string::npos) return "";
    size_t j = s.find_last_not_of(ws);
    return s.substr(i, j - i + 1);
}

std::pair<std::string, std::string> split_filename(const std::string& str) {
    size_t idx = str.find_last_of("/\\");
    if (idx == std::string::npos) return {"./", str};
    return {str.substr(0, idx) + "/", str.substr(idx + 1)};
}

struct LoaderData {
    Scene* s;
    SceneData* sd;
    RenderSettings* rs;
    std::unordered_map<std::string, int> loaded_materials;
};

void parse_camera(LoaderData* ld, std::istringstream& iss) {
    Vec3f o, at, up;
    float tan_fovy;
    if (!(iss >> o.x >> o.y >> o.z >> at.x >> at.y >> at.z >> up.x >> up.y >> up.z >> tan_fovy))
        throw std::runtime_error("Malformed camera line");
    ld->s->camera = std::make_shared<Camera>();
    ld->s->camera->setup(ld->rs->resX, ld->rs->resY, tan_fovy, o, normalized(v(at.x - o.x, at.y - o.y, at.z - o.z)), up);
}

// loader.h:129-141: "light point x y z <colour texture> intensity"; the colour is
// read as an sRGB Vec3 texture from the working directory and evaluated at uv 0
// (PointLight::connect_eye_path, point_light.h:20-28), whose point lookup lands on
// the first texel of an image.  Radiance = colour * intensity / (4 pi), with
// Vec3 / a = Vec3 * float(1.0 / a) and a = float(4 pi) (math.h:104-109).
void parse_light(LoaderData* ld, std::istringstream& iss) {
    std::string type;
    iss >> type;
    if (type != "point") throw std::runtime_error("Unknown light type: " + type);
    Vec3f pos;
    std::string tex;
    float intensity;
    if (!(iss >> pos.x >> pos.y >> pos.z >> tex >> intensity)) throw std::runtime_error("Malformed light line");
    Tex t = get_texture(ld->sd, "", tex, false, true);
    Vec3f col = t.value;
    if (t.image >= 0) {
        const std::vector<float>& px = ld->sd->images[t.image].texels;
        col = v(px[0], px[1], px[2]);
    }
    const float inv = (float)(1.0 / (double)(float)(4.0 * M_PI));
    const float rec[6] = {pos.x, pos.y, pos.z, (col.x * intensity) * inv, (col.y * intensity) * inv,
                          (col.z * intensity) * inv};
    ld->s->point_lights.insert(ld->s->point_lights.end(), rec, rec + 6);
}

void parse_mat(LoaderData* ld, const std::string& dir, const std::string& filename) {
    auto it = ld->loaded_materials.find(filename);
    if (it != ld->loaded_materials.end()) {
        ld->sd->materials.push_back(ld->sd->materials.at(it->second));
        return;
    }
    std::ifstream iss(dir + filename);
    if (!iss) throw std::runtime_error("Cannot open material file: " + dir + filename);
    std::string word;
    iss >> word;
    auto m = std::make_shared<Material>();
    if (word == "DISNEY") {
        std::string albedo, spec, rough, metal, transparent, ior, normal, alpha;
        iss >> albedo >> spec >> rough >> metal >> transparent >> ior >> normal >> alpha;
        if (!iss) throw std::runtime_error("Malformed DISNEY material: " + filename);
        SceneData* sd = ld->sd;
        Tex ta = get_texture(sd, dir, albedo, false, true);
        get_texture(sd, dir, spec, true, true);
        Tex tr = get_texture(sd, dir, rough, true, false);
        Tex tm = get_texture(sd, dir, metal, true, false);
        get_texture(sd, dir, transparent, true, true);
        get_texture(sd, dir, ior, true, true);
        Tex tn = get_texture(sd, dir, normal, false, false);
        get_texture(sd, dir, alpha, true, false);
        // standard_material.h:46-61: metal < 0.1 -> Lambertian(albedo), else GGX(roughness);
        // an image metal texture makes the choice per hit
        if (tm.image >= 0) m->type = CPT_MAT_STANDARD;
        else m->type = tm.value.x < 0.1f ? CPT_MAT_DIFFUSE : CPT_MAT_GLOSSY;
        m->albedo[0] = ta.value.x;
        m->albedo[1] = ta.value.y;
        m->albedo[2] = ta.value.z;
        m->roughness = tr.value.x;
        m->normal_map[0] = tn.value.x;
        m->normal_map[1] = tn.value.y;
        m->normal_map[2] = tn.value.z;
        m->has_normal_map = true;
        m->tex[0] = ta.image;
        m->tex[1] = tr.image;
        m->tex[2] = tm.image;
        m->tex[3] = tn.image;
    } else if (word == "EMISSION") {
        std::string col, inten;
        iss >> col >> inten;
        if (!iss) throw std::runtime_error("Malformed EMISSION material: " + filename);
        Tex tc = get_texture(ld->sd, dir, col, false, true);
        if (tc.image >= 0) throw std::runtime_error("Image emission textures are not supported: " + col);
        Vec3f c = tc.value;
        float intensity = (float)atof(inten.substr(6, inten.length() - 7).c_str());
        m->type = CPT_MAT_EMISSIVE;
        m->emission[0] = c.x * intensity;  // albedo->sample(uv)*intensity (emissive_material.h:28-30)
        m->emission[1] = c.y * intensity;
        m->emission[2] = c.z * intensity;
        m->has_normal_map = false;
    } else if (word == "DIELECTRIC") {
        // addition: refraction_bxdf.h made reachable from scene files
        std::string ior;
        iss >> ior;
        if (!iss) throw std::runtime_error("Malformed DIELECTRIC material: " + filename);
        m->type = CPT_MAT_DIELECTRIC;
        m->ior = get_texture(ld->sd, dir, ior, true, false).value.x;
        m->has_normal_map = false;
    } else {
        throw std::runtime_error("Wrong bsdf type: " + word);
    }
    ld->loaded_materials.insert({filename, (int)ld->sd->materials.size()});
    ld->sd->materials.push_back(m);
}

void parse_obj(LoaderData* ld, const std::string& dir, const std::string& filename) {
    std::ifstream ifs(dir + filename);
    std::string line;
    if (!std::getline(ifs, line)) throw std::runtime_error("Insuficient number of lines when reading file");
    int n_v = 0, n_f = 0;
    std::string type;
    {
        std::istringstream iss(line);
        iss >> n_v >> n_f >> type;
    }
    if (type != "TRIANGLE_MESH") throw std::runtime_error("Unknown mesh type: " + type);
    if (n_v < 0 || n_f < 0) throw std::runtime_error("Negative mesh sizes in " + filename);
    auto mesh = std::make_shared<Mesh>();
    mesh->p.resize(3 * (size_t)n_v);
    mesh->n.resize(3 * (size_t)n_v);
    mesh->b.resize(3 * (size_t)n_v);
    mesh->uv.resize(2 * (size_t)n_v);
    mesh->f.resize(n_f);
    for (int i = 0; i < n_v; i++) {
        if (!std::getline(ifs, line)) throw std::runtime_error("Insuficient number of lines when reading verts");
        std::istringstream iss(line);
        float p[3], uv[2], n[3], t[3];
        iss >> p[0] >> p[1] >> p[2] >> uv[0] >> uv[1] >> n[0] >> n[1] >> n[2] >> t[0] >> t[1] >> t[2];
        Vec3f bt = cross(v(n[0], n[1], n[2]), v(t[0], t[1], t[2]));  // loader.h:237
        for (int k = 0; k < 3; k++) {
            mesh->p[3 * (size_t)i + k] = p[k];
            mesh->n[3 * (size_t)i + k] = n[k];
        }
        mesh->uv[2 * (size_t)i] = uv[0];
        mesh->uv[2 * (size_t)i + 1] = uv[1];
        mesh->b[3 * (size_t)i] = bt.x;
        mesh->b[3 * (size_t)i + 1] = bt.y;
        mesh->b[3 * (size_t)i + 2] = bt.z;
    }
    for (int i = 0; i < n_f; i++) {
        if (!std::getline(ifs, line)) throw std::runtime_error("Insuficient number of lines when reading faces");
        std::istringstream iss(line);
        int a, b, c;
        iss >> a >> b >> c;
        if (a < 0 || b < 0 || c < 0 || a >= n_v || b >= n_v || c >= n_v)
            throw std::runtime_error("Face index out of range in " + filename);
        mesh->f[i] = {a, b, c};
    }
    ld->sd->meshes.push_back(mesh);
}

}  // namespace

void Camera::setup(int resX, int resY, float tf, Vec3f o, Vec3f forward, Vec3f up) {
    // camera_perspective.h:19-27
    image = RgbImage(resX, resY);
    tan_fovy = tf;
    origin = o;
    z = normalized(v(forward.x * (float)(-1.0), forward.y * (float)(-1.0), forward.z * (float)(-1.0)));
    x = normalized(cross(up, z));
    y = normalized(cross(z, x));
    aspect = float(resX) / float(resY);
}

void Scene::flatten(const SceneData& sd) {
    positions.clear();
    normals.clear();
    bitangents.clear();
    indices.clear();
    tri_material.clear();
    light_triangles.clear();
    materials.clear();
    uvs.clear();
    texels.clear();
    textures.clear();
    material_textures.clear();
    if (sd.materials.size() < sd.meshes.size()) throw std::runtime_error("mesh without material");
    // distinct material objects -> table index
    std::vector<const Material*> table;
    std::vector<int> mesh_mat(sd.meshes.size());
    for (size_t i = 0; i < sd.meshes.size(); i++) {
        const Material* m = sd.materials[i].get();
        int id = -1;
        for (size_t k = 0; k < table.size(); k++)
            if (table[k] == m) id = (int)k;
        if (id < 0) {
            id = (int)table.size();
            table.push_back(m);
        }
        mesh_mat[i] = id;
    }
    for (const Material* m : table) {
        float rec[16] = {0};
        rec[0] = (float)m->type;
        rec[1] = m->albedo[0];
        rec[2] = m->albedo[1];
        rec[3] = m->albedo[2];
        rec[4] = m->roughness;
        rec[5] = m->emission[0];
        rec[6] = m->emission[1];
        rec[7] = m->emission[2];
        rec[8] = m->ior;
        rec[9] = m->normal_map[0];
        rec[10] = m->normal_map[1];
        rec[11] = m->normal_map[2];
        rec[12] = m->has_normal_map ? 1.0f : 0.0f;
        materials.insert(materials.end(), rec, rec + 16);
        material_textures.insert(material_textures.end(), m->tex, m->tex + 4);
    }
    if (table.empty()) {
        materials.assign(16, 0.0f);
        material_textures.assign(4, -1);
    }
    for (const TexImage& im : sd.images) {
        textures.push_back(im.w);
        textures.push_back(im.h);
        textures.push_back(im.ch);
        textures.push_back((int32_t)texels.size());
        texels.insert(texels.end(), im.texels.begin(), im.texels.end());
    }
    n_textures = (int)sd.images.size();
    n_materials = (int)std::max<size_t>(1, table.size());
    for (size_t k = 0; k < point_lights.size() / 6; k++) light_triangles.push_back(-1 - (int32_t)k);
    // loader.h:353-367: mesh order, then face order; emissive triangles become lights
    int voff = 0;
    for (size_t i = 0; i < sd.meshes.size(); i++) {
        const Mesh& m = *sd.meshes[i];
        positions.insert(positions.end(), m.p.begin(), m.p.end());
        normals.insert(normals.end(), m.n.begin(), m.n.end());
        bitangents.insert(bitangents.end(), m.b.begin(), m.b.end());
        if (m.uv.size() == m.p.size() / 3 * 2) uvs.insert(uvs.end(), m.uv.begin(), m.uv.end());
        else uvs.insert(uvs.end(), m.p.size() / 3 * 2, 0.0f);
        for (const auto& f : m.f) {
            int tid = (int)tri_material.size();
            indices.push_back(f[0] + voff);
            indices.push_back(f[1] + voff);
            indices.push_back(f[2] + voff);
            tri_material.push_back(mesh_mat[i]);
            if (sd.materials[i]->is_emissive()) light_triangles.push_back(tid);
        }
        voff += (int)(m.p.size() / 3);
    }
    n_vertices = voff;
    n_triangles = (int)tri_material.size();
}

cpt_scene_desc Scene::desc() const {
    cpt_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n_vertices = n_vertices;
    d.positions = positions.data();
    d.normals = normals.data();
    d.bitangents = bitangents.data();
    d.n_triangles = n_triangles;
    d.indices = indices.data();
    d.tri_material = tri_material.data();
    d.n_materials = n_materials;
    d.materials = materials.data();
    d.n_lights = (int32_t)light_triangles.size();
    d.light_triangles = light_triangles.data();
    if (!camera) throw std::runtime_error("scene has no camera");
    const Camera& c = *camera;
    float cam[16] = {c.origin.x, c.origin.y, c.origin.z, c.x.x, c.x.y, c.x.z, c.y.x, c.y.y,
                     c.y.z, c.z.x, c.z.y, c.z.z, c.tan_fovy, c.aspect, (float)c.image.resX, (float)c.image.resY};
    std::memcpy(d.camera, cam, sizeof(cam));
    for (int k = 0; k < 3; k++) d.env[k] = env[k];
    d.uvs = uvs.empty() ? nullptr : uvs.data();
    d.n_textures = n_textures;
    d.textures = textures.empty() ? nullptr : textures.data();
    d.texels = texels.empty() ? nullptr : texels.data();
    d.material_textures = material_textures.data();
    d.n_point_lights = (int32_t)(point_lights.size() / 6);
    d.point_lights = point_lights.empty() ? nullptr : point_lights.data();
    return d;
}

namespace Loader {

void load_render_settings(RenderSettings* rs, const std::string& filename) {
    std::ifstream ifs(filename);
    if (!ifs) throw std::runtime_error("Cannot open settings file: " + filename);
    std::string line;
    while (std::getline(ifs, line)) {
        line = line.substr(0, line.find("#"));
        std::istringstream iss(line);
        std::string type;
        iss >> type;
        if (type == "renderer") {
            std::string rt;
            iss >> rt;
            if (rt == "pathtracer_mis" || rt == "pathtracer_mis_gpu") rs->renderer = RenderSettings::PATHTRACER_MIS_GPU;
            else if (rt == "debug") rs->renderer = RenderSettings::DEBUG;
            else if (rt == "lighttracer") rs->renderer = RenderSettings::LIGHTTRACER;
            else if (rt == "pure_pt") rs->renderer = RenderSettings::PURE_PT;
            else if (rt == "pathtracer") rs->renderer = RenderSettings::PATHTRACER;
            else if (rt == "pathtracer_gpu") rs->renderer = RenderSettings::PATHTRACER_GPU;
            else if (rt == "pure_pt_gpu") rs->renderer = RenderSettings::PURE_PT_GPU;
            else if (rt == "pathtracer_mlt") rs->renderer = RenderSettings::PATHTRACER_MLT;
            else if (rt == "lighttracer_mlt") rs->renderer = RenderSettings::LIGHTTRACER_MLT;
            else if (rt == "photonmapping") rs->renderer = RenderSettings::PHOTONMAPPING;
            else if (rt == "ppg") rs->renderer = RenderSettings::PPG;
            else if (rt == "validation") rs->renderer = RenderSettings::VALIDATION;
            else throw std::runtime_error("Unknown renderer type: " + rt);
        } else if (type == "spp") {
            iss >> rs->spp;
        } else if (type == "resolution") {
            iss >> rs->resX >> rs->resY;
        } else if (type == "scene") {
            iss >> rs->scene_name;
        } else if (type == "output") {
            iss >> rs->output_name;
        } else if (type == "max_depth") {
            iss >> rs->max_depth;
        } else if (type == "seed") {
            iss >> rs->seed;
        } else if (type == "device") {
            iss >> rs->device;
        } else if (type != "") {
            throw std::runtime_error("Unknown render setting: " + type);
        }
    }
}

void load_scene(Scene* s, SceneData* sd, RenderSettings* rs, const std::string& filename) {
    auto dir_file = split_filename(filename);
    LoaderData ld{s, sd, rs, {}};
    std::ifstream ifs(filename);
    if (!ifs) throw std::runtime_error("Cannot open scene file: " + filename);
    std::string line;
    while (std::getline(ifs, line)) {
        line = line.substr(0, line.find("#"));
        line = trim(line);
        std::istringstream iss(line);
        std::string type;
        iss >> type;
        if (type == "camera") {
            parse_camera(&ld, iss);
        } else if (type == "obj") {
            std::string obj, mat;
            iss >> obj >> mat;
            parse_obj(&ld, dir_file.first, obj);
            parse_mat(&ld, dir_file.first, mat);
        } else if (type == "light") {
            parse_light(&ld, iss);
        } else if (type == "environment") {
            // addition: constant environment radiance "environment r g b"
            iss >> s->env[0] >> s->env[1] >> s->env[2];
        } else if (type != "") {
            throw std::runtime_error("Unknown scene obj: " + type);
        }
    }
    if (!s->camera) throw std::runtime_error("Scene has no camera: " + filename);
    s->flatten(*sd);
}

}  // namespace Loader

// ---------------------------------------------------------------- image output
namespace {

uint32_t crc_table[256];
void crc_init() {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}
uint32_t crc(const unsigned char* b, size_t n, uint32_t c = 0xffffffffu) {
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ b[i]) & 0xff] ^ (c >> 8);
    return c;
}
void put32(std::vector<unsigned char>& o, uint32_t x) {
    o.push_back((unsigned char)(x >> 24));
    o.push_back((unsigned char)(x >> 16));
    o.push_back((unsigned char)(x >> 8));
    o.push_back((unsigned char)x);
}
void chunk(FILE* f, const char* type, const std::vector<unsigned char>& data) {
    std::vector<unsigned char> c;
    put32(c, (uint32_t)data.size());
    c.insert(c.end(), type, type + 4);
    c.insert(c.end(), data.begin(), data.end());
    uint32_t cr = crc(c.data() + 4, c.size() - 4) ^ 0xffffffffu;
    put32(c, cr);
    fwrite(c.data(), 1, c.size(), f);
}

// PNG with stored (uncompressed) deflate blocks
void write_png(const std::string& name, int w, int h, const std::vector<unsigned char>& rgb) {
    crc_init();
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) throw std::runtime_error("Cannot write " + name);
    const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    fwrite(sig, 1, 8, f);
    std::vector<unsigned char> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.push_back(8);
    ihdr.push_back(2);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    chunk(f, "IHDR", ihdr);
    std::vector<unsigned char> raw;
    raw.reserve((size_t)h * (3 * w + 1));
    for (int j = 0; j < h; j++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + (size_t)j * 3 * w, rgb.begin() + (size_t)(j + 1) * 3 * w);
    }
    std::vector<unsigned char> z;
    z.push_back(0x78);
    z.push_back(0x01);
    size_t pos = 0;
    do {
        size_t n = std::min<size_t>(65535, raw.size() - pos);
        z.push_back(pos + n == raw.size() ? 1 : 0);
        z.push_back((unsigned char)(n & 0xff));
        z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 0xff));
        z.push_back((unsigned char)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char ch : raw) {
        a = (a + ch) % 65521;
        b = (b + a) % 65521;
    }
    put32(z, (b << 16) | a);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
    fclose(f);
}

}  // namespace

void RgbImage::save(const std::string& filename) const {
    size_t P = (size_t)resX * resY;
    std::vector<unsigned char> rgb8(3 * P);
    for (size_t i = 0; i < 3 * P; i++) rgb8[i] = put2char1(linear2srgb(pixels[i]));
    write_png(filename + ".png", resX, resY, rgb8);
    {  // Radiance RGBE, flat scanlines
        FILE* f = fopen((filename + ".hdr").c_str(), "wb");
        if (!f) throw std::runtime_error("Cannot write " + filename + ".hdr");
        fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", resY, resX);
        std::vector<unsigned char> row(4 * (size_t)resX);
        for (int j = 0; j < resY; j++) {
            for (int i = 0; i < resX; i++) {
                const float* p = &pixels[3 * ((size_t)j * resX + i)];
                float m = std::max(p[0], std::max(p[1], p[2]));
                unsigned char* e = &row[4 * (size_t)i];
                if (!(m >= 1e-32f)) {
                    e[0] = e[1] = e[2] = e[3] = 0;
                } else {
                    int ex;
                    float sc = frexpf(m, &ex) * 256.0f / m;
                    e[0] = (unsigned char)(p[0] * sc);
                    e[1] = (unsigned char)(p[1] * sc);
                    e[2] = (unsigned char)(p[2] * sc);
                    e[3] = (unsigned char)(ex + 128);
                }
            }
            fwrite(row.data(), 1, row.size(), f);
        }
        fclose(f);
    }
    {  // PFM, bottom row first
        FILE* f = fopen((filename + ".pfm").c_str(), "wb");
        if (!f) throw std::runtime_error("Cannot write " + filename + ".pfm");
        fprintf(f, "PF\n%d %d\n-1.0\n", resX, resY);
        for (int j = resY - 1; j >= 0; j--) fwrite(&pixels[3 * (size_t)j * resX], sizeof(float), 3 * (size_t)resX, f);
        fclose(f);
    }
}

}  // namespace cpppt

// ---------------------------------------------------------------- C ABI scene file
// A scene file loaded and flattened behind the C ABI, so that a caller with host
// types of its own (the reference's Scene / SceneData, which share these names)
// can obtain a cpt_scene_desc without linking against this file's C++ classes.
struct cpt_scene_file {
    cpppt::Scene scene;
    cpppt::SceneData data;
    cpppt::RenderSettings settings;
};

static void copy_message(char* err, int32_t err_len, const char* msg) {
    if (!err || err_len <= 0) return;
    std::strncpy(err, msg, (size_t)err_len - 1);
    err[err_len - 1] = 0;
}

extern "C" int cpt_scene_file_load(const char* path, int32_t res_x, int32_t res_y, cpt_scene_file** file,
                                   cpt_scene_desc* desc, char* err, int32_t err_len) {
    copy_message(err, err_len, "");
    if (!path || !file || !desc || res_x <= 0 || res_y <= 0) {
        copy_message(err, err_len, "cpt_scene_file_load: null argument or non-positive resolution");
        return CPT_E_INVALID;
    }
    *file = nullptr;
    cpt_scene_file* f = nullptr;
    try {
        f = new cpt_scene_file();
        f->settings.resX = res_x;
        f->settings.resY = res_y;
        f->settings.scene_name = path;
        cpppt::Loader::load_scene(&f->scene, &f->data, &f->settings, path);
        *desc = f->scene.desc();
    } catch (const std::exception& e) {
        copy_message(err, err_len, e.what());
        delete f;
        return CPT_E_INVALID;
    } catch (...) {
        copy_message(err, err_len, "cpt_scene_file_load: unknown error");
        delete f;
        return CPT_E_INVALID;
    }
    *file = f;
    return CPT_OK;
}

extern "C" void cpt_scene_file_free(cpt_scene_file* file) { delete file; }
