// host_scene.hpp — a scene held on the host before upload: triangle soup, materials,
// optional environment map and camera (what Scene::init receives, Scene.h:84-101),
// plus the albedo textures and per-vertex uv of scenes read from files.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pg_tracer.h"

struct pg_host_scene {
    std::vector<float> pos, nrm;   // 9 floats per triangle each
    std::vector<uint32_t> mat;     // material index per triangle
    std::vector<pg_material> mats;
    std::vector<float> env;        // RGB, env_w x env_h, lat-long y-up
    int32_t env_w = 0, env_h = 0;
    pg_camera_desc cam;
    // textures (scene files only; empty for procedural scenes)
    std::vector<float> uv;         // 6 floats per triangle: u0 v0 u1 v1 u2 v2
    std::vector<int32_t> mat_tex;  // per material: texture index or -1
    std::vector<float> texels;     // RGB of every texture, concatenated
    std::vector<uint64_t> tex_offset;
    std::vector<int32_t> tex_w, tex_h;
    std::vector<std::string> tex_path;
    std::vector<std::string> warnings;
};

namespace pg {
// Largest image the decoders accept (PNG, Radiance HDR, JPEG): 2^28 pixels.
constexpr int64_t kMaxImagePixels = int64_t(1) << 28;
// Decoded image, `channels` values per pixel as the file stores them (stb_image's
// req_comp = 0 convention): 8-bit PNG samples in u8, Radiance HDR floats in f32.
struct Image {
    int32_t w = 0, h = 0, channels = 0;
    std::vector<uint8_t> u8;
    std::vector<float> f32;
};
// Returns an empty string on success, else the reason.
std::string read_png(const std::string& path, Image& out);
std::string read_hdr(const std::string& path, Image& out);
std::string read_jpeg(const std::string& path, Image& out);
// Texture::load's decoding: ".hdr" in the name selects Radiance HDR; otherwise the
// file's signature picks PNG or JPEG (other formats are reported as unsupported).
std::string read_image_file(const std::string& path, Image& out, bool& hdr);
}  // namespace pg
