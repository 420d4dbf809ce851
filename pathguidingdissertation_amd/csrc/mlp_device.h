// mlp_device.h — device functions of the learned guide: input encoding, MFMA
// tiles of the network, and the vMF mixture density and sampling that the render
// kernels use.
//
// MFMA tiles use v_mfma_f32_16x16x32_bf16 (D[16x16] += A[16x32] B[32x16]): lane l
// supplies row l%16 of A and column l%16 of B with eight K values each from lane
// group g = l/16, and receives D rows 4g..4g+3 of column l%16. A wave carries 32
// samples as two 16-column blocks; activations stay in registers between layers.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "mlp.h"

namespace pg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(const u16x8& a, const u16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                   0, 0, 0);
}

// ------------------------------------------------------------------ encoding
__device__ __forceinline__ uint32_t grid_entry(const MlpView& v, int l, uint32_t x, uint32_t y, uint32_t z) {
    if (v.dense[l]) {
        uint32_t n1 = v.res[l] + 1u;
        return x + n1 * (y + n1 * z);
    }
    return ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & v.table_mask;
}

// Cell and fractional position of p01 in [0,1]^3 at level l.
__device__ __forceinline__ void grid_cell(const MlpView& v, int l, const float p01[3], uint32_t ic[3], float fr[3]) {
    const float res = (float)v.res[l];
    for (int a = 0; a < 3; ++a) {
        float p = p01[a] * res;
        int i = (int)floorf(p);
        i = min(max(i, 0), (int)v.res[l] - 1);
        ic[a] = (uint32_t)i;
        fr[a] = p - (float)i;
    }
}

__device__ __forceinline__ float corner_weight(int c, const float fr[3]) {
    float wx = (c & 1) ? fr[0] : 1.f - fr[0];
    float wy = (c & 2) ? fr[1] : 1.f - fr[1];
    float wz = (c & 4) ? fr[2] : 1.f - fr[2];
    return (wx * wy) * wz;
}

// Both features of hash level l at p01 (trilinear over the 8 corners, corner order
// c = x + 2y + 4z).
__device__ __forceinline__ void hash_level(const MlpView& v, int l, const float p01[3], float& f0, float& f1) {
    uint32_t ic[3];
    float fr[3];
    grid_cell(v, l, p01, ic, fr);
    const uint32_t* tab = v.grid + ((size_t)l << v.log2_table);
    float a0 = 0.f, a1 = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        uint32_t e = grid_entry(v, l, ic[0] + (c & 1), ic[1] + ((c >> 1) & 1), ic[2] + (c >> 2));
        uint32_t pr = tab[e];
        float w = corner_weight(c, fr);
        a0 += w * bf2f((uint16_t)(pr & 0xffffu));
        a1 += w * bf2f((uint16_t)(pr >> 16));
    }
    f0 = a0;
    f1 = a1;
}

// One-blob encoding of s in [0,1] with 8 bins (Gaussian kernel of width 1/8).
__device__ __forceinline__ void one_blob8(float s, float out[8]) {
    float t = 8.f * s;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float d = ((float)i + 0.5f) - t;
        out[i] = expf(-0.5f * (d * d));
    }
}

__device__ __forceinline__ void scene_unit(const MlpView& v, F3 x, float p01[3]) {
    p01[0] = fminf(fmaxf((x.x - v.lo[0]) * v.inv_ext, 0.f), 1.f);
    p01[1] = fminf(fmaxf((x.y - v.lo[1]) * v.inv_ext, 0.f), 1.f);
    p01[2] = fminf(fmaxf((x.z - v.lo[2]) * v.inv_ext, 0.f), 1.f);
}

// The 8 inputs 32 kb + 8 g + j (j = 0..7) of one sample as BF16: kb 0 holds hash
// levels 4g..4g+3, kb 1 the blobs of n.x, n.y, n.z, wo.x (g = 0..3), kb 2 the blobs
// of wo.y, wo.z and zeros.
__device__ __forceinline__ void encode_slots(const MlpView& v, bool valid, const float p01[3], F3 n, F3 wo, int g,
                                             u16x8 out[3]) {
    float f[8];
    // hash levels
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float f0 = 0.f, f1 = 0.f;
        if (valid) hash_level(v, 4 * g + q, p01, f0, f1);
        f[2 * q] = f0;
        f[2 * q + 1] = f1;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) out[0][j] = f2bf(f[j]);
    const float comps[6] = {n.x, n.y, n.z, wo.x, wo.y, wo.z};
    // kb 1: component g; kb 2: component 4 + g (g < 2)
    {
        float s = (comps[g] + 1.f) * 0.5f;
        one_blob8(valid ? s : -8.f, f);
#pragma unroll
        for (int j = 0; j < 8; ++j) out[1][j] = valid ? f2bf(f[j]) : (uint16_t)0;
    }
    {
        int ci = 4 + g;
        float s = (g < 2) ? (comps[ci < 6 ? ci : 5] + 1.f) * 0.5f : 0.f;
        one_blob8(s, f);
#pragma unroll
        for (int j = 0; j < 8; ++j) out[2][j] = (valid && g < 2) ? f2bf(f[j]) : (uint16_t)0;
    }
}

// ------------------------------------------------------------------ network tiles
// acc[nb][mb] += W(rows mb*16.., K slots) * in[kb][nb], weights from LDS with row pitch NK*32.
template <int NK, int NM>
__device__ __forceinline__ void layer_fwd(const uint16_t* w, const u16x8 (&in)[NK][2], f32x4 (&acc)[2][NM], int lane) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < NM; ++mb) acc[nb][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mb = 0; mb < NM; ++mb)
#pragma unroll
        for (int kb = 0; kb < NK; ++kb) {
            u16x8 a = *reinterpret_cast<const u16x8*>(w + (mb * 16 + c) * (NK * 32) + kb * 32 + g * 8);
            acc[0][mb] = mfma16(a, in[kb][0], acc[0][mb]);
            acc[1][mb] = mfma16(a, in[kb][1], acc[1][mb]);
        }
}

// bias + ReLU on a 64-row accumulator; packs the BF16 result as the next layer's K
// slots and returns the ReLU mask (bit nb*16 + mb*4 + r).
__device__ __forceinline__ uint32_t bias_relu(const f32x4 (&acc)[2][4], const float* bias, int g, u16x8 (&out)[2][2]) {
    uint32_t mask = 0;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float h = acc[nb][mb][r] + bias[mb * 16 + 4 * g + r];
                bool pos = h > 0.f;
                mask |= (pos ? 1u : 0u) << (nb * 16 + mb * 4 + r);
                out[mb >> 1][nb][((mb & 1) << 2) + r] = f2bf(pos ? h : 0.f);
            }
    return mask;
}

// Backward through one layer: acc[nb][mb] = sum_k W[k][m] d[k][nb] for input rows
// m = mb*16 + lane%16. W rows are the layer's outputs in logical order (n_rows of
// them), columns in MFMA layout with pitch ld; `first` selects the input-layer
// row order 8 (c/4) + 4 mb + c%4 that returns hash features to the lanes that
// computed them.
template <int NM>
__device__ __forceinline__ void layer_bwd(const uint16_t* w, int ld, int n_rows, bool first, const u16x8 (&d)[2][2],
                                          f32x4 (&acc)[2][NM], int lane) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < NM; ++mb) acc[nb][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mb = 0; mb < NM; ++mb) {
        const int col = first ? (8 * (c >> 2) + 4 * mb + (c & 3)) : phys_of(mb * 16 + c);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            u16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                int o = log_of(kb * 32 + 8 * g + j);
                a[j] = (o < n_rows) ? w[o * ld + col] : (uint16_t)0;
            }
            acc[0][mb] = mfma16(a, d[kb][0], acc[0][mb]);
            acc[1][mb] = mfma16(a, d[kb][1], acc[1][mb]);
        }
    }
}

// ------------------------------------------------------------------ vMF mixture
// vMF density kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)): the
// normalisation factor.
__device__ __forceinline__ float vmf_norm(float kappa) {
    return kappa / (6.283185307179586f * (-expm1f(-2.f * kappa)));
}

__device__ __forceinline__ float mix_pdf(const float4* mx, F3 w) {
    const float pt[8] = {mx[10].x, mx[10].y, mx[10].z, mx[10].w, mx[11].x, mx[11].y, mx[11].z, mx[11].w};
    float p = 0.f;
#pragma unroll
    for (int j = 0; j < kLobes; ++j) {
        float4 l = mx[j];
        float t = (l.x * w.x + l.y * w.y) + l.z * w.z;
        p += pt[j] * expf(l.w * (t - 1.f));
    }
    return p;
}

// Sample the mixture: u0 picks the lobe, (u1, u2) the direction (numerically stable
// vMF inversion: cos = 1 + log1p((1 - u1) expm1(-2 kappa)) / kappa).
__device__ __forceinline__ F3 mix_sample(const float4* mx, float u0, float u1, float u2) {
    const float pi_[8] = {mx[8].x, mx[8].y, mx[8].z, mx[8].w, mx[9].x, mx[9].y, mx[9].z, mx[9].w};
    int j = 0;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < kLobes - 1; ++k) {
        acc += pi_[k];
        if (u0 >= acc) j = k + 1;
    }
    float4 l = mx[0];
#pragma unroll
    for (int k = 1; k < kLobes; ++k)
        if (k == j) l = mx[k];
    float kappa = l.w;
    float ct = 1.f + log1pf((1.f - u1) * expm1f(-2.f * kappa)) / kappa;
    ct = fminf(fmaxf(ct, -1.f), 1.f);
    float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
    float ph = 6.283185307179586f * u2;
    Frame fr = frame_from(F3{l.x, l.y, l.z});
    return normalize(to_world(fr, F3{st * cosf(ph), st * sinf(ph), ct}));
}

}  // namespace pg
