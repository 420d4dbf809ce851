// mlp.h — learned directional guide: multiresolution hash-grid + one-blob input
// encoding, a 3 x 64 ReLU network and an 8-lobe von Mises-Fisher mixture with a
// learned BSDF/guide selection probability, trained with Adam on path-vertex records.
//
// The reference has no learned guide; its guiding outline (Renderer.h:398-414) only
// records PathVertex data (Renderer.h:42-47). This module turns those records into
// training samples {x, n, wo, wi, weight, BSDF pdf} and fits q(wi | x, n, wo) to the
// product of incident radiance and BSDF by minimising the radiance-weighted negative
// log-likelihood of the one-sample mixture c q_vmf + (1 - c) p_bsdf.
//
// Encoding (inputs 0..79, 80..95 zero): 16 hash levels x 2 features (resolution
// res_l = floor(res_min * b^l), trilinear, dense index when (res_l+1)^3 fits the
// table, otherwise the xor hash with primes 1, 2654435761, 805459861), one-blob
// encodings of (n+1)/2 and (wo+1)/2 with 8 Gaussian bins each.
// Network: W1 64x96, W2 64x64, W3 64x64, W4 48x64 (41 outputs used); BF16 operands,
// FP32 accumulation (v_mfma_f32_16x16x32_bf16), FP32 biases and master weights.
// Outputs: a[0..7] lobe logits, m[8..31] lobe axes (3 per lobe, normalised),
// k[32..39] log concentrations (clamped to [-6, 9]), s[40] selection logit with
// c = 0.05 + 0.9 sigmoid(s).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pg_common.h"

namespace pg {

constexpr int kHashLevels = 16;
constexpr int kInputs = 96;     // 80 used
constexpr int kInputsUsed = 80;
constexpr int kHidden = 64;
constexpr int kOutputs = 48;    // 41 used
constexpr int kOutputsUsed = 41;
constexpr int kLobes = 8;
constexpr float kKappaLogMin = -6.f;
constexpr float kKappaLogMax = 9.f;
constexpr float kSelMin = 0.05f;
constexpr float kSelRange = 0.9f;

// FP32 master parameters, logical [out][in] row-major layout.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kInputs;    // 6144
constexpr int kOffW2 = kOffB1 + kHidden;              // 6208
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;    // 10304
constexpr int kOffW3 = kOffB2 + kHidden;              // 10368
constexpr int kOffB3 = kOffW3 + kHidden * kHidden;    // 14464
constexpr int kOffW4 = kOffB3 + kHidden;              // 14528
constexpr int kOffB4 = kOffW4 + kOutputs * kHidden;   // 17600
constexpr int kParams = kOffB4 + kOutputs;            // 17648

// BF16 weight mirror in the layout the MFMA tiles read: W1 as is; W2..W4 with
// their input columns permuted so that the K slots of a lane match the rows it
// holds in the previous layer's accumulator (see phys_of / log_of).
constexpr int kMW1 = 0;
constexpr int kMW2 = kMW1 + kHidden * kInputs;   // 6144
constexpr int kMW3 = kMW2 + kHidden * kHidden;   // 10240
constexpr int kMW4 = kMW3 + kHidden * kHidden;   // 14336
constexpr int kMirror = kMW4 + kOutputs * kHidden;  // 17408

// Per-sample BF16 activation row of a training chunk (input of the weight-gradient GEMM).
constexpr int kRowX = 0;
constexpr int kRowH1 = 96;
constexpr int kRowH2 = 160;
constexpr int kRowH3 = 224;
constexpr int kRowD1 = 288;
constexpr int kRowD2 = 352;
constexpr int kRowD3 = 416;
constexpr int kRowDO = 480;
constexpr int kRowLen = 528;

// Mixture of one shading point as the render kernels read it (13 float4):
// [0..7] lobe (mu.xyz, kappa), [8..9] lobe probabilities, [10..11] lobe probability
// times the vMF normalisation kappa / (2 pi (1 - e^{-2 kappa})), [12] (c, 0, 0, 0).
constexpr int kMixF4 = 13;

// Neuron n of a 64-wide layer sits in accumulator block mb = n/16, lane group
// g = (n/4)%4, element r = n%4; as the K operand of the next layer it occupies slot
// phys_of(n) = 32 (mb/2) + 8 g + 4 (mb%2) + r. log_of is the inverse.
PG_HD int phys_of(int n) {
    int mb = n >> 4, g = (n >> 2) & 3, r = n & 3;
    return (mb >> 1) * 32 + g * 8 + ((mb & 1) << 2) + r;
}
PG_HD int log_of(int p) {
    int kb = p >> 5, g = (p >> 3) & 3, j = p & 7;
    return 16 * (2 * kb + (j >> 2)) + 4 * g + (j & 3);
}

// Round-to-nearest-even FP32 -> BF16 (NaN stays NaN).
PG_HD uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
PG_HD float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Read-only view of the network for the encode/forward kernels.
struct MlpView {
    const uint16_t* wmirror;  // kMirror BF16 weights (MFMA layout)
    const float* params;      // FP32 master parameters (biases read from here)
    const uint32_t* grid;     // [level][table] two BF16 features per 32-bit entry
    uint32_t table_mask;      // entries per level - 1 (power of two)
    uint32_t log2_table;
    uint32_t res[kHashLevels];
    uint32_t dense[kHashLevels];
    float lo[3];
    float inv_ext;            // position -> [0,1]^3 of the scene cube
    // hash-table gradient accumulators: level l keeps gcopies[l] private copies of
    // its gents[l] entries (two floats each) from float offset goff[l]; a block adds
    // into copy blockIdx % gcopies[l]. Copies scale with the expected hits per entry
    // (about 2^20 / entries, 1..64), so coarse levels, which every sample touches,
    // are not serialised on a few thousand addresses, while the whole accumulator
    // (~120 MB at 2^19 entries per level) stays cache-resident
    uint32_t gcopies[kHashLevels];
    uint32_t gents[kHashLevels];
    uint64_t goff[kHashLevels];
};
constexpr uint32_t kGradCopiesMax = 64;
constexpr uint32_t kGradHitsPerCopy = 1u << 20;

// One training sample (64 bytes): position + normalised weight, normal + BSDF pdf
// of wi, outgoing direction, incident (sampled) direction.
struct alignas(16) TrainSample {
    float4 x_w;
    float4 n_pb;
    float4 wo;
    float4 wi;
};

// Training data that the record pass keeps beside each 48-byte record (32 bytes):
// a = (wo.xyz, lum(Li * f |cos| / (q rr))), b = (BSDF pdf of wi, sampling pdf q, 0, 0).
struct alignas(16) TrainAux {
    float4 a;
    float4 b;
};

struct AdamParams {
    float lr, beta1, beta2, eps;
    float bc1, bc2;  // 1 - beta^t
};

struct Record;
hipError_t launch_mlp_gather(const Record* rec, const TrainAux* aux, uint32_t n_rec, uint32_t offset, uint32_t stride,
                             uint32_t batch, float w_scale, float w_clip, TrainSample* out, hipStream_t stream);
// Key/value arrays of n entries for a radix sort, with its histogram and scan scratch
// (sort_tiles(n) * 256 and scan_tiles(sort_tiles(n) * 256) + 1 entries).
struct SortBuffers {
    uint32_t *keys, *vals, *keys_tmp, *vals_tmp, *hist, *scan;
};
// Reorders n training samples by the Morton code of their positions (in -> out).
hipError_t launch_mlp_sort_batch(const MlpView& v, const TrainSample* in, uint32_t n, const SortBuffers& sb,
                                 TrainSample* out, hipStream_t stream);
hipError_t launch_mlp_train_chunk(const MlpView& v, const TrainSample* smp, uint32_t n, float inv_batch,
                                  uint16_t* rows, float* loss, float* grid_grad, int nan_inject, int cu_count,
                                  hipStream_t stream);
uint32_t mlp_wgrad_groups(uint32_t n);
hipError_t launch_mlp_wgrad(const uint16_t* rows, uint32_t n, uint32_t groups, float* partial, int accumulate,
                            hipStream_t stream);
hipError_t launch_mlp_adam(float* params, float* m, float* v, uint16_t* mirror, const float* partial, uint32_t groups,
                           const AdamParams& ap, int update, float* grad_out, uint32_t* status, hipStream_t stream);
hipError_t launch_grid_adam(const MlpView& view, float* grid_master, uint32_t* grid_bf, float* m, float* v, float* grad,
                            size_t n, const AdamParams& ap, int update, float* grad_out, uint32_t* status,
                            hipStream_t stream);
// Deterministic FP64 sum of in[i * stride], i < n; part: kSumParts doubles of scratch.
constexpr int kSumParts = 256;
hipError_t launch_sum(const float* in, size_t n, size_t stride, double* part, float* out, hipStream_t stream);
hipError_t launch_mlp_mirror(const float* params, uint16_t* mirror, hipStream_t stream);
// Level resolutions, dense/hashed indexing and the scene cube of a view (host).
void mlp_view_frame(MlpView& v, uint32_t log2_table, float res_min, float res_max, const float* centre, float radius);
hipError_t launch_grid_mirror(const float* grid_master, uint32_t* grid_bf, size_t n_entries, hipStream_t stream);
hipError_t launch_mlp_eval(const MlpView& v, const float4* x, const float4* nrm, const float4* wo, uint32_t n,
                           float4* mix, int cu_count, hipStream_t stream);
hipError_t launch_mix_query(const float4* mix, uint32_t n, const float4* wi_eval, const float4* u, float* pdf_eval,
                            float4* wi_out, hipStream_t stream);

}  // namespace pg
