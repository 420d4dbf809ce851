// pgd_render.cpp — example host program over the C++ mirror (pgd.hpp): the reference's
// render loop (main.cpp:147-199) with the GPU path behind the same class names.
//
//   pgd_render <config> <spp> <seed> <out.bin> [--learned [train_every steps
//              [max_depth env_visibility [record_samples]]]]
//                                                 render a procedural scene (unguided, or
//                                                 with the learned guide trained online:
//                                                 a guided record pass + `steps` Adam steps
//                                                 every `train_every` SPP, default 1 and 1,
//                                                 guided passes between; max depth default
//                                                 8, environment shadow rays along the light
//                                                 direction when env_visibility is 1, record
//                                                 passes of record_samples samples, default 1);
//                                                 writes int32 width, height, spp + film sum
//   pgd_render --scene <dir> <spp> <seed> <out.png> [out.hdr]
//                                                 the reference's main.cpp flow on one of its
//                                                 scene directories: loadScene, render, savePNG
//                                                 (and saveHDR)
//   pgd_render --selftest                        host-only checks (no GPU needed)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/pgd/pgd.hpp"

static int selftest() {
    // Metrics.h:8-69 on a known case: errors 1, -1, 2 (and 0s) over a 2x1 image
    std::vector<float> truth = {1.f, 2.f, 3.f, 4.f, 5.f, 6.f};
    std::vector<float> test = {2.f, 1.f, 5.f, 4.f, 5.f, 6.f};
    pgd::Metrics m;
    float mse = m.calculateMSE(test, truth, 2, 1);
    float rmse = m.calculateRMSE(test, truth, 2, 1);
    float rrmse = m.calculateRRMSE(test, truth, 2, 1);
    if (mse != 1.f || rmse != 1.f || std::fabs(rrmse - 1.f / (std::sqrt(91.f) + 1e-2f)) > 1e-7f) {
        std::printf("metrics mismatch %g %g %g\n", mse, rmse, rrmse);
        return 1;
    }
    std::vector<float> short_truth(truth.begin(), truth.end() - 1);
    if (m.calculateMSE(test, short_truth, 2, 1) != -1.f) return 1;
    pgd::SceneData c = pgd::SceneData::generate(1, 0);
    if (c.triangles() != 36 || c.camera.width != 256.f) {
        std::printf("cornell box: %u triangles\n", c.triangles());
        return 1;
    }
    std::printf("selftest ok\n");
    return 0;
}

static int render_scene_dir(int argc, char** argv) {
    try {
        pgd::SceneData data = pgd::SceneData::load(argv[2]);
        for (const std::string& w : data.warnings) std::printf("warning: %s\n", w.c_str());
        pgd::Device dev(0);
        pgd::Scene scene(dev, data);
        pg_render_params p = pg_render_params_default();
        p.seed = (uint32_t)std::atoi(argv[4]);
        pgd::RayTracer rt(scene, p);
        rt.clear();
        const int spp = std::atoi(argv[3]);
        for (int i = 0; i < spp; ++i) rt.render();
        rt.savePNG(argv[5]);
        if (argc == 7) rt.saveHDR(argv[6]);
        std::printf("rendered %s: %u triangles, %zu textures, %d spp on %s\n", argv[2], data.triangles(),
                    data.tex_width.size(), rt.getSPP(), dev.name().c_str());
    } catch (const pgd::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 4;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "--selftest") == 0) return selftest();
    if ((argc == 6 || argc == 7) && std::strcmp(argv[1], "--scene") == 0) return render_scene_dir(argc, argv);
    const bool learned =
        (argc == 6 || argc == 8 || argc == 10 || argc == 11) && std::strcmp(argv[5], "--learned") == 0;
    if (argc != 5 && !learned) {
        std::fprintf(stderr,
                     "usage: %s <config> <spp> <seed> <out.bin> [--learned [train_every steps [max_depth "
                     "env_visibility [record_samples]]]] | --selftest\n",
                     argv[0]);
        return 2;
    }
    pg_mlp_schedule schedule{1, 1, 1, 0, 0, 1, {0, 0}};
    if (argc >= 8) {
        schedule.train_every = (uint32_t)std::atoi(argv[6]);
        schedule.steps = (uint32_t)std::atoi(argv[7]);
    }
    if (argc == 11) schedule.record_samples = (uint32_t)std::atoi(argv[10]);
    try {
        pgd::SceneData data = pgd::SceneData::generate(std::atoi(argv[1]), 0);
        pgd::Device dev(0);
        pgd::Scene scene(dev, data);
        pg_render_params p = pg_render_params_default();
        p.seed = (uint32_t)std::atoi(argv[3]);
        if (argc >= 10) {
            p.max_depth = (uint32_t)std::atoi(argv[8]);
            p.env_visibility = (uint32_t)std::atoi(argv[9]);
        }
        pgd::RayTracer rt(scene, p);
        rt.clear();
        const int spp = std::atoi(argv[2]);
        if (learned) {
            p.record_nee = 1;
            pgd::LearnedGuide guide(dev);
            guide.render_online(p, (uint32_t)spp, schedule);
            const pg_mlp_info info = guide.info();
            std::printf("learned guide: %u Adam steps, last loss %g%s\n", info.step, info.last_loss,
                        info.diverged ? " (diverged)" : "");
        } else {
            for (int i = 0; i < spp; ++i) rt.render();  // one sample per pixel per call, as the reference loop
        }
        uint32_t n = 0;
        std::vector<float> film = rt.film(&n);
        FILE* f = std::fopen(argv[4], "wb");
        if (!f) return 3;
        int32_t hdr[3] = {(int32_t)data.camera.width, (int32_t)data.camera.height, (int32_t)n};
        std::fwrite(hdr, 4, 3, f);
        std::fwrite(film.data(), 4, film.size(), f);
        std::fclose(f);
        std::printf("rendered %d spp on %s\n", rt.getSPP(), dev.name().c_str());
    } catch (const pgd::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 4;
    }
    return 0;
}
