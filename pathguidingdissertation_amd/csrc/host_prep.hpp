// host_prep.hpp — host-side scene preparation for the device path: reference-order
// BVH build, camera matrices, environment tables and device record packing.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pg_tracer.h"
#include "pg_common.h"

namespace pg {

struct HV3 {
    float x = 0.f, y = 0.f, z = 0.f;
    HV3() {}
    HV3(float a, float b, float c) : x(a), y(b), z(c) {}
    HV3 operator+(const HV3& v) const { return HV3(x + v.x, y + v.y, z + v.z); }
    HV3 operator-(const HV3& v) const { return HV3(x - v.x, y - v.y, z - v.z); }
    HV3 operator*(float s) const { return HV3(x * s, y * s, z * s); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

// Binned-SAH BVH in the reference's topology and triangle order (Geometry.h:293-446).
struct HostBvh {
    std::vector<BvhNode> nodes;      // reference numbering (root 0, children adjacent)
    std::vector<int32_t> tri_index;  // BVHNode::triangleIndexes
    uint32_t max_depth = 0;
    double build_ms = 0.0;
};
void build_bvh(const float* positions, uint32_t n_tris, HostBvh& out);

// Camera matrices as SceneLoader.h:226-244 and Scene.h:25-43 compute them.
CameraRec make_camera(const pg_camera_desc& d);

// Environment tables: the reference Distribution2D's texel pdf (Sampling.h:166-209)
// and a Vose alias table over the same discrete distribution.
struct EnvTables {
    int32_t w = 0, h = 0;
    std::vector<float> texels;   // RGBA32F, A unused
    std::vector<float> pdf;      // f[v][u] / marginal integral (uv-space density)
    std::vector<float> prob;     // alias acceptance probability
    std::vector<uint32_t> alias; // alias target
    float marginal_integral = 0.f;
    float power = 0.f;           // EnvironmentMap::totalIntegratedPower (Lights.h:200-210)
};
void build_env_tables(const float* rgb, int32_t w, int32_t h, EnvTables& out);

// Everything the device needs, packed (see pg_common.h for the layouts).
struct PreparedScene {
    HostBvh bvh;
    std::vector<BvhNode> node_slots;  // nodes at slot index + 1
    std::vector<TriIsect> tris;       // leaf order
    std::vector<TriShade> shade;      // original order
    std::vector<MaterialRec> mats;
    std::vector<LightRec> lights;
    EnvTables env;
    bool has_env = false;
    CameraRec camera;
    float centre[3] = {0, 0, 0};
    float radius = 0.f;
};
std::string prepare_scene(const float* positions, const float* normals, const uint32_t* mat_idx,
                          uint32_t n_tris, const pg_material* mats, uint32_t n_mats,
                          const float* env_rgb, int32_t env_w, int32_t env_h,
                          const pg_camera_desc& cam, PreparedScene& out);

}  // namespace pg
