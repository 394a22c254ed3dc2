// cpt_render: command-line entry with the reference's conventions
// (src/main.cpp:89-146): argv[1] = settings file, renderer chosen by the
// settings token, wall time of render() printed as "<seconds>s", runtime
// errors printed.  Renderer tokens `pathtracer_mis` / `pathtracer_mis_gpu`,
// `pathtracer` / `pathtracer_gpu`, `pure_pt` / `pure_pt_gpu`, `ppg` and `lighttracer` select the
// GPU integrators; the reference's other algorithms are not part of this build
// and are reported as such.
#include <chrono>
#include <iostream>
#include <memory>

#include "cpt_host.h"

using namespace cpppt;

int main(int argc, char** argv) {
    try {
        if (argc < 2) {
            std::cout << "Insuficient Arguments" << std::endl;
            return 0;
        }
        RenderSettings rs;
        Scene s;
        SceneData sd;
        Loader::load_render_settings(&rs, argv[1]);
        Loader::load_scene(&s, &sd, &rs, rs.scene_name);
        std::unique_ptr<Renderer> renderer;
        if (rs.renderer == RenderSettings::PATHTRACER_MIS_GPU) {
            renderer = std::make_unique<PathtracerMISGPU>(rs);
        } else if (rs.renderer == RenderSettings::PATHTRACER || rs.renderer == RenderSettings::PATHTRACER_GPU) {
            renderer = std::make_unique<PathtracerGPU>(rs);
        } else if (rs.renderer == RenderSettings::PURE_PT || rs.renderer == RenderSettings::PURE_PT_GPU) {
            renderer = std::make_unique<PurePtGPU>(rs);
        } else if (rs.renderer == RenderSettings::PPG) {
            renderer = std::make_unique<PpgGPU>(rs);
        } else if (rs.renderer == RenderSettings::LIGHTTRACER) {
            renderer = std::make_unique<LighttracerGPU>(rs);
        } else {
            throw std::runtime_error(
                "renderer not provided by this build (pathtracer_mis, pathtracer, pure_pt, ppg, lighttracer)");
        }
        auto t1 = std::chrono::high_resolution_clock::now();
        renderer->render(s, rs.output_name);
        auto t2 = std::chrono::high_resolution_clock::now();
        std::chrono::duration<double, std::milli> ms = t2 - t1;
        std::cout << ms.count() * 0.001 << "s\n";
    } catch (const std::runtime_error& e) {
        std::cout << e.what() << std::endl;
        return 1;
    }
    return 0;
}
