// PathtracerMISGPU: the Renderer plug-in that replaces PathtracerMIS
// (src/renderer/pathtracerMIS.h:133-167) with the HIP wavefront integrator.
#include <iostream>

#include "cpt_host.h"

namespace cpppt {

namespace {
void check(cpt_ctx* ctx, int rc, const char* what) {
    if (rc != CPT_OK) {
        std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
        if (ctx) msg += ": " + std::string(cpt_last_error(ctx));
        throw std::runtime_error(msg);  // the reference's error convention (main.cpp:143-145)
    }
}
}  // namespace

PathtracerMISGPU::PathtracerMISGPU(const RenderSettings& rs, int integrator_flags)
    : rs_(rs), flags_(integrator_flags) {
    int rc = cpt_create(&ctx_, rs.device);
    if (rc != CPT_OK) throw std::runtime_error("cpt_create failed: no usable gfx950 device " + std::to_string(rs.device));
}

PathtracerMISGPU::~PathtracerMISGPU() { cpt_destroy(ctx_); }

void PathtracerMISGPU::render(Scene& sc, std::string filename) {
    RgbImage& image = sc.camera->get_image();
    cpt_scene_desc d = sc.desc();
    check(ctx_, cpt_upload_scene(ctx_, &d), "cpt_upload_scene");
    cpt_render_params p{};
    p.width = image.resX;
    p.height = image.resY;
    p.spp_axis = rs_.spp;
    p.max_depth = rs_.max_depth;
    p.seed = rs_.seed;
    p.flags = flags_;
    check(ctx_, cpt_render(ctx_, &p, &stats_), "cpt_render");
    std::vector<float> rgb(3 * (size_t)image.resX * image.resY);
    check(ctx_, cpt_read_film(ctx_, rgb.data(), nullptr, nullptr), "cpt_read_film");
    for (int j = 0; j < image.resY; j++)
        for (int i = 0; i < image.resX; i++) {
            const float* q = &rgb[3 * ((size_t)i + (size_t)j * image.resX)];
            Vec3f px;
            px.x = q[0];
            px.y = q[1];
            px.z = q[2];
            image.put_pixel(i, j, px);
        }
    image.save(filename);
}

}  // namespace cpppt
