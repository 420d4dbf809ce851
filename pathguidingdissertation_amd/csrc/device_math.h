// device_math.h — FP32 vector arithmetic for the gfx950 kernels in the
// reference's operation order (MyMath.h:75-449). Built with -ffp-contract=off so
// no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>

#include "pg_common.h"

namespace pg {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float a, float b, float c) { return F3{a, b, c}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return F3{(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}
__device__ __forceinline__ float length(F3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float length_sq(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ F3 normalize(F3 a) {
    float l = 1.f / length(a);
    return F3{a.x * l, a.y * l, a.z * l};
}
__device__ __forceinline__ F3 ld3(const float* p) { return F3{p[0], p[1], p[2]}; }

// Colour (MyMath.h:30-73): division by a scalar multiplies by its reciprocal.
struct C3 {
    float r, g, b;
};
__device__ __forceinline__ C3 c3(float a, float b, float c) { return C3{a, b, c}; }
__device__ __forceinline__ C3 operator+(C3 a, C3 b) { return C3{a.r + b.r, a.g + b.g, a.b + b.b}; }
__device__ __forceinline__ C3 operator-(C3 a, C3 b) { return C3{a.r - b.r, a.g - b.g, a.b - b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, C3 b) { return C3{a.r * b.r, a.g * b.g, a.b * b.b}; }
__device__ __forceinline__ C3 operator/(C3 a, C3 b) { return C3{a.r / b.r, a.g / b.g, a.b / b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, float s) { return C3{a.r * s, a.g * s, a.b * s}; }
__device__ __forceinline__ C3 operator/(C3 a, float s) {
    float iv = 1.f / s;
    return C3{a.r * iv, a.g * iv, a.b * iv};
}
__device__ __forceinline__ float lum(C3 c) { return ((0.2126f * c.r) + (0.7152f * c.g)) + (0.0722f * c.b); }
__device__ __forceinline__ bool finite3(C3 c) {
    return !(isnan(c.r) || isnan(c.g) || isnan(c.b) || isinf(c.r) || isinf(c.g) || isinf(c.b));
}
__device__ __forceinline__ bool is_zero(C3 c) { return c.r == 0.f && c.g == 0.f && c.b == 0.f; }

// Orthonormal frame (MyMath.h:406-433)
struct Frame {
    F3 u, v, w;
};
__device__ __forceinline__ Frame frame_from(F3 n) {
    Frame f;
    f.w = normalize(n);
    if (fabsf(f.w.x) > fabsf(f.w.y)) {
        float l = 1.f / sqrtf(f.w.x * f.w.x + f.w.z * f.w.z);
        f.u = F3{f.w.z * l, 0.f, -f.w.x * l};
    } else {
        float l = 1.f / sqrtf(f.w.y * f.w.y + f.w.z * f.w.z);
        f.u = F3{0.f, f.w.z * l, -f.w.y * l};
    }
    f.v = cross(f.w, f.u);
    return f;
}
__device__ __forceinline__ F3 to_local(const Frame& f, F3 a) { return F3{dot(a, f.u), dot(a, f.v), dot(a, f.w)}; }
__device__ __forceinline__ F3 to_world(const Frame& f, F3 a) { return (f.u * a.x + f.v * a.y) + f.w * a.z; }

__device__ __forceinline__ F3 spherical_to_world(float theta, float phi) {
    return F3{cosf(phi) * sinf(theta), sinf(phi) * sinf(theta), cosf(theta)};
}

__device__ __forceinline__ float power_heuristic(float pa, float pb) {  // MyMath.h:23-28
    float a = pa * pa, b = pb * pb;
    if (a + b <= 0.f) return 0.f;
    return a / (a + b);
}

// Block-aggregated appends to two counters (both queues a shading lane may extend):
// one atomic per counter per block of BLOCK lanes, issued together by thread 0, so
// the device sees one atomic per counter per 256 lanes (device-scope atomics on one
// address are serialised at memory, so their count bounds a shading kernel). Slots follow
// lane order within a wave and wave order within the block. Every thread of the block
// must call it (it holds barriers); lds: 2 * (BLOCK / 64) + 2 words.
template <int BLOCK>
__device__ __forceinline__ void block_append2(uint32_t* ca, bool wa, uint32_t& sa, uint32_t* cb, bool wb,
                                              uint32_t& sb, uint32_t* lds) {
    constexpr int kWaves = BLOCK / 64;
    const unsigned long long ma = __ballot(wa), mb = __ballot(wb);
    const int lane = __lane_id(), w = (int)(threadIdx.x >> 6);
    if (lane == 0) {
        lds[w] = (uint32_t)__popcll(ma);
        lds[kWaves + w] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tb = 0;
        for (int k = 0; k < kWaves; ++k) {
            const uint32_t x = lds[k], y = lds[kWaves + k];
            lds[k] = ta;
            lds[kWaves + k] = tb;
            ta += x;
            tb += y;
        }
        const uint32_t ba = ta ? atomicAdd(ca, ta) : 0u;
        const uint32_t bb = tb ? atomicAdd(cb, tb) : 0u;
        lds[2 * kWaves] = ba;
        lds[2 * kWaves + 1] = bb;
    }
    __syncthreads();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    sa = lds[2 * kWaves] + lds[w] + (uint32_t)__popcll(ma & below);
    sb = lds[2 * kWaves + 1] + lds[kWaves + w] + (uint32_t)__popcll(mb & below);
    __syncthreads();  // lds is rewritten by the next call
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* counter, unsigned long long v) {
    // reduce across the active lanes, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    unsigned long long act = __ballot(1);
    int leader = __ffsll((long long)act) - 1;
    if (__lane_id() == leader && v) atomicAdd(counter, v);
}

}  // namespace pg
