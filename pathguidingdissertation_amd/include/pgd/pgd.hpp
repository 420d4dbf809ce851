// pgd.hpp — header-only C++ mirror of the reference's hot-path classes over the C ABI
// of include/pg_tracer.h (library libpgtrace.so).
//
// The reference's render loop uses Scene (Scene.h:74-176), RayTracer
// (Renderer.h:236-751) and Metrics (Metrics.h:6-70) directly; these classes keep
// those names and methods so that the loop can switch to the GPU path by changing
// types. Errors from the C ABI become pgd::Error exceptions carrying pg_status.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/pg_tracer.h"

namespace pgd {

class Error : public std::runtime_error {
public:
    Error(pg_status code, const std::string& what) : std::runtime_error(what), code_(code) {}
    pg_status code() const { return code_; }

private:
    pg_status code_;
};

inline void check(pg_status s, const pg_ctx* ctx = nullptr) {
    if (s != PG_OK) {
        const char* m = pg_last_error(ctx);
        throw Error(s, std::string("pg error ") + std::to_string(s) + ": " + (m ? m : ""));
    }
}

// One GPU context (RayTracer::init's device side, Renderer.h:251-263).
class Device {
public:
    explicit Device(int index = 0) { check(pg_create(&ctx_, index)); }
    ~Device() { if (ctx_) pg_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    pg_ctx* get() const { return ctx_; }
    std::string name() const {
        char buf[128] = {0};
        int32_t cu = 0;
        double peak = 0.0;
        check(pg_device_info(ctx_, buf, sizeof(buf), &cu, &peak), ctx_);
        return buf;
    }

private:
    pg_ctx* ctx_ = nullptr;
};

// Triangle soup + materials + environment + camera, as Scene::init receives them.
struct SceneData {
    std::vector<float> positions;     // 9 floats per triangle (v0, v1, v2)
    std::vector<float> normals;       // 9 floats per triangle
    std::vector<uint32_t> material_index;
    std::vector<pg_material> materials;
    std::vector<float> env_rgb;       // lat-long RGB32F (empty: no environment map)
    int32_t env_width = 0, env_height = 0;
    pg_camera_desc camera{};
    // albedo textures of scenes read from files (empty for procedural scenes)
    std::vector<float> uv;                 // 6 floats per triangle
    std::vector<int32_t> material_texture; // per material, -1 = constant albedo
    std::vector<float> texels;             // RGB of every texture, concatenated
    std::vector<uint64_t> texel_offset;
    std::vector<int32_t> tex_width, tex_height;
    std::vector<std::string> warnings;
    uint32_t triangles() const { return (uint32_t)material_index.size(); }
    bool textured() const { return !tex_width.empty(); }

    // loadScene (SceneLoader.h:185-269): a reference scene directory.
    static SceneData load(const std::string& dir, uint32_t flags = 0) {
        pg_host_scene* h = nullptr;
        pg_status st = pg_scene_load(dir.c_str(), flags, &h);
        if (st != PG_OK) throw Error(st, pg_scene_load_error());
        SceneData d = copy_out(h);
        uint32_t ntex = 0, nwarn = 0;
        uint64_t ntexel = 0;
        st = pg_scene_texture_sizes(h, &ntex, &ntexel, &nwarn);
        if (st == PG_OK) {
            d.uv.resize((size_t)d.triangles() * 6);
            d.material_texture.resize(d.materials.size());
            d.texels.resize((size_t)ntexel * 3);
            d.texel_offset.resize(ntex);
            d.tex_width.resize(ntex);
            d.tex_height.resize(ntex);
            st = pg_scene_copy_textures(h, d.uv.data(), d.material_texture.data(), d.texels.data(),
                                        d.texel_offset.data(), d.tex_width.data(), d.tex_height.data());
            for (uint32_t i = 0; i < nwarn; ++i) d.warnings.push_back(pg_scene_warning(h, i));
        }
        pg_scene_free(h);
        check(st);
        return d;
    }

    // Procedural scenes of the benchmark configurations (pg_scene_generate).
    static SceneData generate(int32_t config, uint32_t seed) {
        pg_host_scene* h = nullptr;
        check(pg_scene_generate(config, seed, &h));
        SceneData d;
        try {
            d = copy_out(h);
        } catch (...) {
            pg_scene_free(h);
            throw;
        }
        pg_scene_free(h);
        return d;
    }

private:
    static SceneData copy_out(pg_host_scene* h) {
        SceneData d;
        uint32_t nt = 0, nm = 0;
        int32_t ew = 0, eh = 0;
        pg_status s = pg_scene_sizes(h, &nt, &nm, &ew, &eh);
        if (s == PG_OK) {
            d.positions.resize((size_t)nt * 9);
            d.normals.resize((size_t)nt * 9);
            d.material_index.resize(nt);
            d.materials.resize(nm);
            d.env_rgb.resize((size_t)ew * eh * 3);
            d.env_width = ew;
            d.env_height = eh;
            s = pg_scene_copy(h, d.positions.data(), d.normals.data(), d.material_index.data(), d.materials.data(),
                              d.env_rgb.empty() ? nullptr : d.env_rgb.data(), &d.camera);
        }
        check(s);
        return d;
    }
};

// Scene::init + Scene::build (Scene.h:84-137); traverse / visible (Scene.h:103-146).
class Scene {
public:
    Scene(Device& dev, const SceneData& d) : dev_(&dev), data_(&d) {
        check(pg_scene_upload(dev.get(), d.positions.data(), d.normals.data(), d.material_index.data(),
                              d.triangles(), d.materials.data(), (uint32_t)d.materials.size(),
                              d.env_rgb.empty() ? nullptr : d.env_rgb.data(), d.env_width, d.env_height,
                              &d.camera),
              dev.get());
        if (d.textured())
            check(pg_scene_textures(dev.get(), d.uv.data(), d.material_texture.data(), (uint32_t)d.materials.size(),
                                    d.texels.data(), d.texel_offset.data(), d.tex_width.data(), d.tex_height.data(),
                                    (uint32_t)d.tex_width.size()),
                  dev.get());
    }
    std::vector<pg_hit> traverse(const std::vector<pg_ray>& rays) const {
        std::vector<pg_hit> hits(rays.size());
        check(pg_trace_closest(dev_->get(), rays.data(), (uint32_t)rays.size(), hits.data()), dev_->get());
        return hits;
    }
    std::vector<uint8_t> visible(const std::vector<pg_ray>& rays, const std::vector<float>& max_t) const {
        std::vector<uint8_t> vis(rays.size());
        check(pg_trace_shadow(dev_->get(), rays.data(), max_t.data(), (uint32_t)rays.size(), vis.data()),
              dev_->get());
        return vis;
    }
    Device& device() const { return *dev_; }
    const SceneData& data() const { return *data_; }

private:
    Device* dev_;
    const SceneData* data_;
};

// RayTracer (Renderer.h:236-751): clear / render (one sample per pixel) / getSPP /
// the film divided by SPP (Film::save, Imaging.h:263-272).
class RayTracer {
public:
    explicit RayTracer(Scene& scene, pg_render_params params = pg_render_params_default())
        : scene_(&scene), params_(params) {}
    pg_render_params& params() { return params_; }
    void clear() { check(pg_film_clear(ctx()), ctx()); }
    void render(uint32_t spp = 1) { check(pg_render_spp(ctx(), &params_, spp), ctx()); }
    int getSPP() const {  // NOLINT(readability-identifier-naming): reference name
        uint32_t spp = 0;
        check(pg_film_read(ctx(), nullptr, &spp), ctx());
        return (int)spp;
    }
    std::vector<float> film(uint32_t* spp = nullptr) const {
        const pg_camera_desc& c = scene_->data().camera;
        std::vector<float> f((size_t)c.width * (size_t)c.height * 3);
        check(pg_film_read(ctx(), f.data(), spp), ctx());
        return f;
    }
    std::vector<float> image() const {
        uint32_t spp = 0;
        std::vector<float> f = film(&spp);
        if (spp)
            for (float& v : f) v /= (float)spp;
        return f;
    }
    pg_stats stats() const {
        pg_stats s{};
        check(pg_stats_read(ctx(), &s), ctx());
        return s;
    }
    // Film::tonemap of every pixel (Imaging.h:220-244), 8-bit RGB.
    std::vector<uint8_t> tonemap(float exposure = 1.f, float lmax_white = 10.f) const {
        const pg_camera_desc& c = scene_->data().camera;
        std::vector<uint8_t> out((size_t)c.width * (size_t)c.height * 3);
        check(pg_film_tonemap(ctx(), exposure, lmax_white, out.data()), ctx());
        return out;
    }
    void savePNG(const std::string& filename) const {  // NOLINT: reference name (Renderer.h:751)
        const pg_camera_desc& c = scene_->data().camera;
        check(pg_image_write_png(filename.c_str(), tonemap().data(), (int32_t)c.width, (int32_t)c.height));
    }
    void saveHDR(const std::string& filename) const {  // NOLINT: reference name (Renderer.h:750)
        const pg_camera_desc& c = scene_->data().camera;
        check(pg_image_write_hdr(filename.c_str(), image().data(), (int32_t)c.width, (int32_t)c.height));
    }

private:
    pg_ctx* ctx() const { return scene_->device().get(); }
    Scene* scene_;
    pg_render_params params_;
};

// Metrics (Metrics.h:6-70): same signatures; -1 on size mismatch as the reference.
struct Metrics {
    static pg_image_metrics all(const std::vector<float>& test, const std::vector<float>& truth, int width,
                                int height) {
        pg_image_metrics m{};
        pg_metrics(test.data(), truth.data(), test.size(), truth.size(), width, height, &m);
        return m;
    }
    float calculateMSE(std::vector<float>& t, std::vector<float>& g, int w, int h) { return all(t, g, w, h).mse; }
    float calculateRMSE(std::vector<float>& t, std::vector<float>& g, int w, int h) { return all(t, g, w, h).rmse; }
    float calculateRRMSE(std::vector<float>& t, std::vector<float>& g, int w, int h) { return all(t, g, w, h).rrmse; }
};

// Learned guide (no reference counterpart): init, train on recorded vertices, render
// with online training.
class LearnedGuide {
public:
    explicit LearnedGuide(Device& dev, pg_mlp_config cfg = pg_mlp_config_default()) : dev_(&dev) {
        check(pg_mlp_init(dev.get(), &cfg), dev.get());
    }
    std::vector<float> train(uint32_t steps) {
        std::vector<float> losses(steps);
        check(pg_mlp_train(dev_->get(), steps, losses.data()), dev_->get());
        return losses;
    }
    void render_online(pg_render_params params, uint32_t spp, pg_mlp_schedule schedule = {1, 1, 1, 0, 0, 1, {0, 0}}) {
        check(pg_mlp_render_online(dev_->get(), &params, spp, &schedule), dev_->get());
    }
    pg_mlp_info info() const {
        pg_mlp_info i{};
        check(pg_mlp_info_read(dev_->get(), &i), dev_->get());
        return i;
    }
    // FP32 network parameters and hash grid (for pg_batch_set_guide or a checkpoint).
    std::pair<std::vector<float>, std::vector<float>> params() const {
        const pg_mlp_info i = info();
        std::vector<float> p(i.n_params), g(i.n_grid_params);
        check(pg_mlp_get_params(dev_->get(), p.data(), g.data()), dev_->get());
        return {std::move(p), std::move(g)};
    }

private:
    Device* dev_;
};

// Batches of views (no reference counterpart).
class Batch {
public:
    explicit Batch(Device& dev) : dev_(&dev) { check(pg_batch_create(dev.get(), &b_), dev.get()); }
    ~Batch() { if (b_) pg_batch_destroy(b_); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    uint32_t add_scene(const SceneData& d) {
        uint32_t id = 0;
        check(pg_batch_add_scene(b_, d.positions.data(), d.normals.data(), d.material_index.data(), d.triangles(),
                                 d.materials.data(), (uint32_t)d.materials.size(),
                                 d.env_rgb.empty() ? nullptr : d.env_rgb.data(), d.env_width, d.env_height, &id),
              dev_->get());
        return id;
    }
    uint32_t add_view(uint32_t scene, const pg_camera_desc& cam) {
        uint32_t id = 0;
        check(pg_batch_add_view(b_, scene, &cam, &id), dev_->get());
        views_.push_back(cam);
        return id;
    }
    std::pair<pg_batch_config, double> autotune(pg_render_params p = pg_render_params_default()) {
        pg_batch_config c{};
        double ms = 0.0;
        check(pg_batch_autotune(b_, &p, &c, &ms), dev_->get());
        return {c, ms};
    }
    // Per-scene learned guide; render with params.integrator = PG_INTEGRATOR_GUIDED_MLP.
    void set_guide(uint32_t scene, const pg_mlp_config& cfg, const std::vector<float>& params,
                   const std::vector<float>& grid) {
        check(pg_batch_set_guide(b_, scene, &cfg, params.data(), grid.data()), dev_->get());
    }
    void clear() { check(pg_batch_clear(b_), dev_->get()); }
    void render(uint32_t spp, pg_render_params p = pg_render_params_default()) {
        check(pg_batch_render(b_, &p, spp), dev_->get());
    }
    std::vector<float> film(uint32_t view, uint32_t* spp = nullptr) const {
        const pg_camera_desc& c = views_.at(view);
        std::vector<float> f((size_t)c.width * (size_t)c.height * 3);
        check(pg_batch_film_read(b_, view, f.data(), spp), dev_->get());
        return f;
    }

private:
    Device* dev_;
    pg_batch* b_ = nullptr;
    std::vector<pg_camera_desc> views_;
};

}  // namespace pgd
