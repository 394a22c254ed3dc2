// Host-side mirror of cpp-pt's plugin interface for the PT-NEE-MIS path.
//
// Same names, argument meaning and error behaviour as the reference:
//   RenderSettings      src/renderer/renderer.h:11-28 (+ max_depth, seed, device)
//   Renderer::render    src/renderer/renderer.h:30-33
//   Loader::load_render_settings / load_scene   src/loader/loader.h:256-374
//   Scene / SceneData   src/scene.h:16-26
//   CameraPerspective   src/camera/camera_perspective.h:19-27
//   RgbImage::put_pixel / save   src/image/rgb_image.h:208-210, 300-318
// Errors are std::runtime_error, as in the reference (loader.h:94,140,203,...).
// The scene is kept flattened (SoA arrays in load order) because that is what
// crosses the C ABI; mesh/material/light order follows loader.h:353-367.
#pragma once
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "cpt.h"

namespace cpppt {

struct Vec3f {
    float x = 0, y = 0, z = 0;
};

struct RenderSettings {
    enum RendererType {
        PATHTRACER,
        PATHTRACER_MIS,
        LIGHTTRACER,
        DEBUG,
        PATHTRACER_MLT,
        LIGHTTRACER_MLT,
        PHOTONMAPPING,
        PURE_PT,
        PPG,
        VALIDATION,
        PATHTRACER_MIS_GPU,
        PATHTRACER_GPU,  // Pathtracer (pathtracer.h) on the GPU kernels
        PURE_PT_GPU      // PurePt (pure_pt.h) on the GPU kernels
    };
    RendererType renderer = RendererType::DEBUG;
    int spp = 1;  // per axis: spp*spp samples per pixel
    int resX = 256;
    int resY = 256;
    std::string scene_name = "";
    std::string output_name = "";
    // additions of this implementation
    int max_depth = 32;     // the reference's hard cap (pathtracerMIS.h:36)
    uint64_t seed = 1;
    int device = 0;
};

// Film with the reference's layout: pixels[i + j*resX], j = 0 at the top.
class RgbImage {
public:
    explicit RgbImage(int w = 0, int h = 0) : resX(w), resY(h), pixels(3 * (size_t)w * h, 0.0f) {}
    void put_pixel(int i, int j, const Vec3f& p) {
        size_t k = 3 * ((size_t)i + (size_t)j * resX);
        pixels.at(k) = p.x;
        pixels.at(k + 1) = p.y;
        pixels.at(k + 2) = p.z;
    }
    // <filename>.png (sRGB 8-bit) and <filename>.hdr (RGBE) like rgb_image.h:300-318,
    // plus <filename>.pfm (lossless float32)
    void save(const std::string& filename) const;
    int resX, resY;
    std::vector<float> pixels;
};

struct Camera {
    Vec3f origin, x, y, z;  // frame columns
    float tan_fovy = 0.35411f;
    float aspect = 1.0f;
    RgbImage image;
    // CameraPerspective(resolution, tan_fovy, origin, forward, up)
    void setup(int resX, int resY, float tan_fovy, Vec3f origin, Vec3f forward, Vec3f up);
    RgbImage& get_image() { return image; }
};

// Image<T> (rgb_image.h:161-197): linear floats, row 0 at the top, 3 channels
// (Vec3 textures) or 1 (float textures)
struct TexImage {
    int w = 0, h = 0, ch = 0;
    std::vector<float> texels;
};

// load_png: 8-bit samples with the file's channel count (stbi_load(..., 0) semantics)
void load_png(const std::string& path, int* w, int* h, int* n, std::vector<uint8_t>* out);

struct Material {
    int type = CPT_MAT_DIFFUSE;
    int tex[4] = {-1, -1, -1, -1};  // albedo, roughness, metal, normal-map images (SceneData::images)
    float albedo[3] = {0, 0, 0};
    float roughness = 0;
    float emission[3] = {0, 0, 0};
    float ior = 1.5f;
    float normal_map[3] = {0.5f, 0.5f, 1.0f};
    bool has_normal_map = true;
    bool is_emissive() const { return type == CPT_MAT_EMISSIVE; }
};

struct Mesh {
    std::vector<float> p, n, b;          // per vertex xyz (b = bitangent)
    std::vector<float> uv;               // per vertex texture coordinates
    std::vector<std::array<int, 3>> f;   // faces
};

struct SceneData {
    std::vector<std::shared_ptr<Mesh>> meshes;
    std::vector<std::shared_ptr<Material>> materials;  // one entry per mesh, load order
    std::vector<TexImage> images;                      // image textures
    std::unordered_map<std::string, int> image_index;  // (file, colour space, bw) -> images
};

struct Scene {
    std::shared_ptr<Camera> camera;
    float env[3] = {0, 0, 0};
    // flattened
    std::vector<float> positions, normals, bitangents;
    std::vector<int32_t> indices, tri_material, light_triangles;
    std::vector<float> materials;  // 16 floats per material
    std::vector<float> uvs, texels;
    std::vector<int32_t> textures, material_textures;
    // point lights in .ptscene order: position, radiance (6 floats each); they come
    // first in the LightGroup (loader.h:129-141 runs while parsing, the shape lights
    // are added after it, loader.h:353-367)
    std::vector<float> point_lights;
    int n_textures = 0;
    int n_vertices = 0, n_triangles = 0, n_materials = 0;
    void flatten(const SceneData& sd);
    cpt_scene_desc desc() const;
};

namespace Loader {
void load_render_settings(RenderSettings* rs, const std::string& filename);
void load_scene(Scene* s, SceneData* sd, RenderSettings* rs, const std::string& filename);
}  // namespace Loader

class Renderer {
public:
    virtual void render(Scene& sc, std::string filename) = 0;
    virtual ~Renderer() {}
};

// GPU drop-in for PathtracerMIS (pathtracerMIS.h:17-168).
class PathtracerMISGPU : public Renderer {
public:
    // integrator_flags: 0 (PathtracerMIS), CPT_RENDER_NEE_ONLY or CPT_RENDER_PURE_PT
    explicit PathtracerMISGPU(const RenderSettings& rs, int integrator_flags = 0);
    ~PathtracerMISGPU() override;
    void render(Scene& sc, std::string filename) override;
    const cpt_render_stats& last_stats() const { return stats_; }

private:
    RenderSettings rs_;
    int flags_;
    cpt_ctx* ctx_ = nullptr;
    cpt_render_stats stats_{};
};

// The reference's Pathtracer (src/renderer/pathtracer.h:15-137): NEE without MIS.
class PathtracerGPU : public PathtracerMISGPU {
public:
    explicit PathtracerGPU(const RenderSettings& rs) : PathtracerMISGPU(rs, CPT_RENDER_NEE_ONLY) {}
};

// The reference's PurePt (src/renderer/pure_pt.h:15-106): BSDF sampling only.
class PurePtGPU : public PathtracerMISGPU {
public:
    explicit PurePtGPU(const RenderSettings& rs) : PathtracerMISGPU(rs, CPT_RENDER_PURE_PT) {}
};

// The reference's Lighttracer (src/renderer/lighttracer.h:15-154): paths from the
// lights, every non-delta vertex connected to the camera.
class LighttracerGPU : public PathtracerMISGPU {
public:
    explicit LighttracerGPU(const RenderSettings& rs) : PathtracerMISGPU(rs, CPT_RENDER_LIGHT_TRACE) {}
};

// The reference's PPG (src/renderer/ppg.h:17-214): SD-tree guided, iterations of
// 1, 2, 4, ... spp, the image of the last one.
class PpgGPU : public PathtracerMISGPU {
public:
    explicit PpgGPU(const RenderSettings& rs) : PathtracerMISGPU(rs, CPT_RENDER_PPG) {}
};

}  // namespace cpppt
