// host_prep.hpp — host-side scene preparation for the device path: reference-order
// BVH build, camera matrices, environment tables and device record packing.
#pragma once
#include <cstdint>
#include <cstring>
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

// 4x4 matrix in the reference's row-major layout and FP32 operation order
// (MyMath.h:151-404).
struct M4 {
    float m[16];
    M4() { std::memset(m, 0, sizeof(m)); m[0] = m[5] = m[10] = m[15] = 1.f; }
    HV3 mul_vec(const HV3& v) const {                                           // MyMath.h:253-259
        return HV3((v.x * m[0] + v.y * m[1]) + v.z * m[2], (v.x * m[4] + v.y * m[5]) + v.z * m[6],
                   (v.x * m[8] + v.y * m[9]) + v.z * m[10]);
    }
    M4 transpose() const {                                                      // MyMath.h:187-194
        M4 r;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) r.m[4 * i + j] = m[4 * j + i];
        return r;
    }
    HV3 mul_point(const HV3& v) const {
        return HV3(((v.x * m[0] + v.y * m[1]) + v.z * m[2]) + m[3],
                   ((v.x * m[4] + v.y * m[5]) + v.z * m[6]) + m[7],
                   ((v.x * m[8] + v.y * m[9]) + v.z * m[10]) + m[11]);
    }
    // Cofactor inverse in the reference's term order (MyMath.h:286-323): per entry,
    // six signed triple products accumulated left to right.
    M4 inverse() const {
        static const signed char T[16][6][4] = {
            {{1,5,10,15},{-1,5,11,14},{-1,9,6,15},{1,9,7,14},{1,13,6,11},{-1,13,7,10}},
            {{-1,1,10,15},{1,1,11,14},{1,9,2,15},{-1,9,3,14},{-1,13,2,11},{1,13,3,10}},
            {{1,1,6,15},{-1,1,7,14},{-1,5,2,15},{1,5,3,14},{1,13,2,7},{-1,13,3,6}},
            {{-1,1,6,11},{1,1,7,10},{1,5,2,11},{-1,5,3,10},{-1,9,2,7},{1,9,3,6}},
            {{-1,4,10,15},{1,4,11,14},{1,8,6,15},{-1,8,7,14},{-1,12,6,11},{1,12,7,10}},
            {{1,0,10,15},{-1,0,11,14},{-1,8,2,15},{1,8,3,14},{1,12,2,11},{-1,12,3,10}},
            {{-1,0,6,15},{1,0,7,14},{1,4,2,15},{-1,4,3,14},{-1,12,2,7},{1,12,3,6}},
            {{1,0,6,11},{-1,0,7,10},{-1,4,2,11},{1,4,3,10},{1,8,2,7},{-1,8,3,6}},
            {{1,4,9,15},{-1,4,11,13},{-1,8,5,15},{1,8,7,13},{1,12,5,11},{-1,12,7,9}},
            {{-1,0,9,15},{1,0,11,13},{1,8,1,15},{-1,8,3,13},{-1,12,1,11},{1,12,3,9}},
            {{1,0,5,15},{-1,0,7,13},{-1,4,1,15},{1,4,3,13},{1,12,1,7},{-1,12,3,5}},
            {{-1,0,5,11},{1,0,7,9},{1,4,1,11},{-1,4,3,9},{-1,8,1,7},{1,8,3,5}},
            {{-1,4,9,14},{1,4,10,13},{1,8,5,14},{-1,8,6,13},{-1,12,5,10},{1,12,6,9}},
            {{1,0,9,14},{-1,0,10,13},{-1,8,1,14},{1,8,2,13},{1,12,1,10},{-1,12,2,9}},
            {{-1,0,5,14},{1,0,6,13},{1,4,1,14},{-1,4,2,13},{-1,12,1,6},{1,12,2,5}},
            {{1,0,5,10},{-1,0,6,9},{-1,4,1,10},{1,4,2,9},{1,8,1,6},{-1,8,2,5}},
        };
        M4 r;
        for (int e = 0; e < 16; ++e) {
            float acc = 0.f;
            for (int t = 0; t < 6; ++t) {
                const signed char* q = T[e][t];
                float p = (m[q[1]] * m[q[2]]) * m[q[3]];
                acc = (t == 0) ? (q[0] > 0 ? p : -p) : (q[0] > 0 ? acc + p : acc - p);
            }
            r.m[e] = acc;
        }
        float det = ((m[0] * r.m[0] + m[1] * r.m[4]) + m[2] * r.m[8]) + m[3] * r.m[12];
        if (det == 0.f) det = 1.f;
        det = 1.f / det;
        for (int i = 0; i < 16; ++i) r.m[i] = r.m[i] * det;
        return r;
    }
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
    bool finite_bounds = true;  // every node bound finite (traversal may use IEEE min/max)
};
// Device material record of a pg_material (false: unknown type).
bool material_record(const pg_material& m, MaterialRec& r);
std::string prepare_scene(const float* positions, const float* normals, const uint32_t* mat_idx,
                          uint32_t n_tris, const pg_material* mats, uint32_t n_mats,
                          const float* env_rgb, int32_t env_w, int32_t env_h,
                          const pg_camera_desc& cam, PreparedScene& out, HostBvh* prebuilt = nullptr);

}  // namespace pg
