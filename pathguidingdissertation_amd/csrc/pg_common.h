// pg_common.h — data layouts and arithmetic shared by the host preparation code
// (g++) and the gfx950 kernels (hipcc). Everything here follows the reference's
// FP32 operation order so that host-built structures and device results agree
// with the reference bit for bit where the reference's arithmetic allows it.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PG_HD __host__ __device__ __forceinline__
#else
#define PG_HD inline
#endif

namespace pg {

// Reference constants (Geometry.h:13-20, Renderer.h:561-568).
constexpr float kEps = 1e-4f;
constexpr float kMtEps = 1e-8f;
constexpr int kLeafTris = 4;
constexpr float kBoundsCost = 1.f;
constexpr float kIsectCost = 1.5f;
constexpr int kBins = 16;
constexpr float kFltMax = 3.402823466e+38f;
constexpr uint32_t kMiss = 0xffffffffu;

// ------------------------------------------------------------------ device layout
// BVH node, 32 bytes. Interior: left_first = index of the left child, the right
// child is left_first + 1 (Geometry.h:255-261). Leaf: count > 0 triangles starting
// at left_first in leaf order. Stored at slot (index + 1) so that every child pair
// (2k+1, 2k+2) starts on a 64-byte boundary.
struct alignas(16) BvhNode {
    float bmin[3];
    int32_t left_first;
    float bmax[3];
    int32_t count;
};
static_assert(sizeof(BvhNode) == 32, "32-byte BVH node");

// Leaf-ordered triangle, 48 bytes: v0 + original ID, e0 = v1 - v0, e1 = v2 - v0
// (exactly the edges Triangle::rayIntersect recomputes, Geometry.h:103-104).
struct alignas(16) TriIsect {
    float v0[3];
    uint32_t id;
    float e0[3];
    float pad0;
    float e1[3];
    float pad1;
};
static_assert(sizeof(TriIsect) == 48, "48-byte triangle");

// Shading attributes by original triangle ID, 64 bytes.
struct alignas(16) TriShade {
    float n0[3];
    uint32_t mat;
    float n1[3];
    float area;
    float n2[3];
    float pad;
    float gn[3];  // Triangle::gNormal() (Geometry.h:153)
    float pad2;
};
static_assert(sizeof(TriShade) == 64, "64-byte shading record");

// Light list entry (Scene.h:93-100, 121-137): kind 0 area triangle, 1 env map.
struct alignas(16) LightRec {
    float p0[3];
    int32_t kind;
    float p1[3];
    float area;
    float p2[3];
    int32_t tri;
    float emission[3];
    float pad;
    float gn[3];
    float pad2;
};
static_assert(sizeof(LightRec) == 80, "80-byte light");

struct alignas(16) MaterialRec {
    int32_t type;
    float albedo[3];
    float emission[3];
    float alpha;  // 1.62142 * sqrt(roughness)
    float eta[3];
    float int_ior;
    float k[3];
    float ext_ior;
    uint32_t is_light;
    uint32_t specular;
    uint32_t two_sided;
    int32_t tex;  // albedo texture (-1: the constant albedo above)
};
static_assert(sizeof(MaterialRec) == 80, "80-byte material");

struct CameraRec {
    float inv_proj[16];
    float cam[16];
    float origin[3];
    float width, height;
};

// ------------------------------------------------------------------ counter RNG
// Keyed hash RNG: one draw = one bijective 32-bit mix of (path key + dim * phi).
// The CPU oracle implements the identical function so both sides draw the same
// numbers for (seed, sample, pixel, dim).
PG_HD uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
PG_HD uint32_t path_key(uint32_t seed, uint32_t sample, uint32_t pixel) {
    return hash32(pixel + hash32(sample + hash32(seed)));
}
PG_HD float keyed_rand(uint32_t key, uint32_t dim) {
    return (float)(hash32(key + dim * 0x9E3779B9u) >> 8) * (1.0f / 16777216.0f);
}

// ------------------------------------------------------------------ float helpers
// std::min / std::max semantics (MyMath.h:143-144), including NaN propagation.
PG_HD float smin(float a, float b) { return (b < a) ? b : a; }
PG_HD float smax(float a, float b) { return (a < b) ? b : a; }

}  // namespace pg
