// mlp.hip — gfx950 kernels of the learned guide.
//
// Training step over a batch (one or more chunks under the activation budget):
//   k_mlp_gather   records -> 64-byte training samples (strided pseudo-random pick)
//   k_mlp_train    per 32-sample wave tile: encode -> 4 BF16 MFMA layers -> vMF
//                  mixture NLL and its output gradient -> 4 MFMA backward layers
//                  -> hash-grid gradient scatter (FP32 atomics); writes each
//                  sample's BF16 activations and deltas (one 1056-byte row)
//   k_mlp_wgrad    weight gradients as a split-K MFMA GEMM over the rows: each
//                  group owns a fixed sample slice and writes one partial vector
//   k_mlp_adam     fixed-order sum of the partials + Adam + BF16 mirror
//   k_grid_adam    Adam over the hash table (+ BF16 mirror, gradient reset)
// Rendering: k_mlp_infer evaluates the network for every shading point of a bounce
// and stores its mixture (13 float4) by queue slot for the shade kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device_math.h"
#include "guide.h"
#include "guiding.h"
#include "kernels.h"
#include "mlp_device.h"

namespace pg {

constexpr int kMlpBlock = 256;    // 4 waves, 128 samples per block tile
constexpr int kStagePitch = 49;   // floats per sample in the per-wave output staging
constexpr int kBiasCount = 3 * kHidden + kOutputs;  // 240

__device__ __forceinline__ void load_net(const MlpView& v, uint16_t* sw, float* sb) {
    const uint4* src = reinterpret_cast<const uint4*>(v.wmirror);
    uint4* dst = reinterpret_cast<uint4*>(sw);
    for (int i = threadIdx.x; i < kMirror / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kBiasCount; i += blockDim.x) {
        int off = i < 64 ? kOffB1 + i : i < 128 ? kOffB2 + (i - 64) : i < 192 ? kOffB3 + (i - 128) : kOffB4 + (i - 192);
        sb[i] = v.params[off];
    }
    __syncthreads();
}

struct FwdOut {
    uint32_t m1, m2, m3;
    f32x4 o[2][3];
};

// Forward of one wave tile; `rows` (nullable) receives the BF16 activations of
// sample s at rows + s * kRowLen for samples below n.
__device__ __forceinline__ void tile_forward(const uint16_t* sw, const float* sb, const u16x8 (&x)[3][2], FwdOut& f,
                                             uint16_t* rows, uint32_t base, uint32_t n, int lane) {
    const int c = lane & 15, g = lane >> 4;
    f32x4 acc[2][4];
    u16x8 h[2][2], h2[2][2];
    auto store_h = [&](const u16x8 (&hh)[2][2], int off) {
        if (!rows) return;
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            uint32_t s = base + nb * 16 + c;
            if (s >= n) continue;
            uint16_t* r = rows + (size_t)s * kRowLen + off;
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const u16x8& p = hh[mb >> 1][nb];
                int q = (mb & 1) << 2;
                uint2 val = make_uint2((uint32_t)p[q] | ((uint32_t)p[q + 1] << 16),
                                       (uint32_t)p[q + 2] | ((uint32_t)p[q + 3] << 16));
                *reinterpret_cast<uint2*>(r + mb * 16 + 4 * g) = val;
            }
        }
    };
    layer_fwd<3, 4>(sw + kMW1, x, acc, lane);
    f.m1 = bias_relu(acc, sb, g, h);
    store_h(h, kRowH1);
    layer_fwd<2, 4>(sw + kMW2, h, acc, lane);
    f.m2 = bias_relu(acc, sb + 64, g, h2);
    store_h(h2, kRowH2);
    layer_fwd<2, 4>(sw + kMW3, h2, acc, lane);
    f.m3 = bias_relu(acc, sb + 128, g, h);
    store_h(h, kRowH3);
    layer_fwd<2, 3>(sw + kMW4, h, f.o, lane);
}

// Output logits of the tile -> per-wave staging [32][kStagePitch] (bias added).
__device__ __forceinline__ void stage_outputs(const FwdOut& f, const float* sb, float* st, int lane) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < 3; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int o = mb * 16 + 4 * g + r;
                st[(nb * 16 + c) * kStagePitch + o] = f.o[nb][mb][r] + sb[192 + o];
            }
}

// Half nb (samples nb*16 .. nb*16+15) of the tile's output logits -> per-wave staging
// [16][kStagePitch] (bias added): half the LDS of stage_outputs.
__device__ __forceinline__ void stage_half(const FwdOut& f, const float* sb, float* st, int lane, int nb) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < 3; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int o = mb * 16 + 4 * g + r;
            st[c * kStagePitch + o] = (nb == 0 ? f.o[0][mb][r] : f.o[1][mb][r]) + sb[192 + o];
        }
}

// Mixture parameters of one sample's outputs (13 float4); c = 0 marks an unusable
// output (non-finite values).
__device__ __forceinline__ void head_mixture(const float* o, float4* mx) {
    float amax = o[0];
#pragma unroll
    for (int j = 1; j < kLobes; ++j) amax = fmaxf(amax, o[j]);
    float e[kLobes], Z = 0.f;
#pragma unroll
    for (int j = 0; j < kLobes; ++j) {
        e[j] = expf(o[j] - amax);
        Z += e[j];
    }
    float pi_[kLobes], pt[kLobes];
    bool ok = isfinite(Z) && Z > 0.f;
#pragma unroll
    for (int j = 0; j < kLobes; ++j) {
        float mxv = o[8 + 3 * j], myv = o[9 + 3 * j], mzv = o[10 + 3 * j];
        float ln = sqrtf((mxv * mxv + myv * myv) + mzv * mzv);
        F3 mu = ln > 1e-20f ? F3{mxv / ln, myv / ln, mzv / ln} : F3{0.f, 0.f, 1.f};
        float kc = fminf(fmaxf(o[32 + j], kKappaLogMin), kKappaLogMax);
        float kappa = expf(kc);
        pi_[j] = e[j] / Z;
        pt[j] = pi_[j] * vmf_norm(kappa);
        ok = ok && isfinite(mu.x) && isfinite(mu.y) && isfinite(mu.z) && isfinite(kappa) && isfinite(pt[j]);
        mx[j] = make_float4(mu.x, mu.y, mu.z, kappa);
    }
    float sg = 1.f / (1.f + expf(-o[40]));
    float cc = kSelMin + kSelRange * sg;
    ok = ok && isfinite(cc);
    mx[8] = make_float4(pi_[0], pi_[1], pi_[2], pi_[3]);
    mx[9] = make_float4(pi_[4], pi_[5], pi_[6], pi_[7]);
    mx[10] = make_float4(pt[0], pt[1], pt[2], pt[3]);
    mx[11] = make_float4(pt[4], pt[5], pt[6], pt[7]);
    mx[12] = make_float4(ok ? cc : 0.f, 0.f, 0.f, 0.f);
}

// Loss -w log(c q_vmf(wi) + (1 - c) p_bsdf) * inv_b of one sample and its gradient
// with respect to the 48 outputs (written over o). Responsibilities are formed in
// the log domain so that sharp lobes cannot overflow.
__device__ __forceinline__ float head_loss_grad(float* o, F3 wi, float pb, float wgt, float inv_b) {
    if (!(wgt > 0.f)) {
        for (int k = 0; k < kOutputs; ++k) o[k] = 0.f;
        return 0.f;
    }
    float amax = o[0];
#pragma unroll
    for (int j = 1; j < kLobes; ++j) amax = fmaxf(amax, o[j]);
    float e[kLobes], Z = 0.f;
#pragma unroll
    for (int j = 0; j < kLobes; ++j) {
        e[j] = expf(o[j] - amax);
        Z += e[j];
    }
    const float logZ = logf(Z);
    float lq[kLobes], kap[kLobes], td[kLobes], len[kLobes];
    F3 mu[kLobes];
    bool clamped[kLobes];
    float M = -INFINITY;
#pragma unroll
    for (int j = 0; j < kLobes; ++j) {
        float mxv = o[8 + 3 * j], myv = o[9 + 3 * j], mzv = o[10 + 3 * j];
        float ln = sqrtf((mxv * mxv + myv * myv) + mzv * mzv);
        len[j] = ln;
        mu[j] = ln > 1e-20f ? F3{mxv / ln, myv / ln, mzv / ln} : F3{0.f, 0.f, 1.f};
        float kr = o[32 + j];
        float kc = fminf(fmaxf(kr, kKappaLogMin), kKappaLogMax);
        clamped[j] = kc != kr;
        float k = expf(kc);
        kap[j] = k;
        float t = (mu[j].x * wi.x + mu[j].y * wi.y) + mu[j].z * wi.z;
        td[j] = t;
        float lognorm = kc - 1.8378770664093453f - logf(-expm1f(-2.f * k));  // log(kappa / (2 pi (1 - e^-2k)))
        lq[j] = ((o[j] - amax) - logZ) + lognorm + k * (t - 1.f);
        M = fmaxf(M, lq[j]);
    }
    float S = 0.f;
#pragma unroll
    for (int j = 0; j < kLobes; ++j) S += expf(lq[j] - M);
    const float sg = 1.f / (1.f + expf(-o[40]));
    const float cc = kSelMin + kSelRange * sg;
    const float la = logf(cc) + M + logf(S);
    const float lb = pb > 0.f ? logf(1.f - cc) + logf(pb) : -INFINITY;
    const float mxl = fmaxf(la, lb);
    const float logP = mxl + logf(expf(la - mxl) + expf(lb - mxl));
    const float G = wgt * inv_b;
    const float rv = expf(la - logP);
    const float rb = pb > 0.f ? expf(lb - logP) : 0.f;
    const float lc = logf(cc);
#pragma unroll
    for (int j = 0; j < kLobes; ++j) {
        float pij = e[j] / Z;
        float rj = expf((lc + lq[j]) - logP);
        o[j] = -G * (rj - pij * rv);
        float k = kap[j];
        o[32 + j] = clamped[j] ? 0.f : -G * rj * ((1.f - (2.f * k) / expm1f(2.f * k)) + k * (td[j] - 1.f));
        float coef = len[j] > 1e-20f ? (-G * rj * k) / len[j] : 0.f;
        o[8 + 3 * j] = coef * (wi.x - td[j] * mu[j].x);
        o[9 + 3 * j] = coef * (wi.y - td[j] * mu[j].y);
        o[10 + 3 * j] = coef * (wi.z - td[j] * mu[j].z);
    }
    o[40] = -G * ((rv / cc) - (rb / (1.f - cc))) * (kSelRange * sg * (1.f - sg));
    for (int k = kOutputsUsed; k < kOutputs; ++k) o[k] = 0.f;
    return -wgt * logP * inv_b;
}

// DPP moves inside the 16-lane rows of a wave: lane c of a row reads lane c - D
// (row_shr) or c + D (row_shl); a lane whose source falls outside its row keeps `fill`.
template <int D>
__device__ __forceinline__ uint32_t row_shr(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x110 + D, 0xf, 0xf, false);
}
template <int D>
__device__ __forceinline__ uint32_t row_shl(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x100 + D, 0xf, 0xf, false);
}
template <int D>
__device__ __forceinline__ void run_add(int c, uint32_t start, float& a, float& b) {
    const float au = __uint_as_float(row_shr<D>(__float_as_uint(a), 0u));
    const float bu = __uint_as_float(row_shr<D>(__float_as_uint(b), 0u));
    if (c - D >= (int)start) {
        a += au;
        b += bu;
    }
}

// Segmented sum over runs of equal keys among consecutive lanes of a 16-lane row
// (lane c = position in the row). Every lane must execute it. Returns true on the
// last lane of each run, which then holds the run's totals in (a, b). Keys must not
// equal 0xFFFFFFFE.
__device__ __forceinline__ bool row_run_sum(uint32_t key, int c, float& a, float& b) {
    constexpr uint32_t kOutside = 0xFFFFFFFEu;
    const bool head = row_shr<1>(key, kOutside) != key;
    uint32_t start = head ? (uint32_t)c : 0u;
    start = max(start, row_shr<1>(start, 0u));
    start = max(start, row_shr<2>(start, 0u));
    start = max(start, row_shr<4>(start, 0u));
    start = max(start, row_shr<8>(start, 0u));
    run_add<1>(c, start, a, b);
    run_add<2>(c, start, a, b);
    run_add<4>(c, start, a, b);
    run_add<8>(c, start, a, b);
    return row_shl<1>(key, kOutside) != key;
}

// Apply a ReLU mask to a 64-row delta accumulator, pack it as the next backward
// layer's K slots and store it (BF16) in the activation rows.
__device__ __forceinline__ void mask_pack(const f32x4 (&acc)[2][4], uint32_t mask, u16x8 (&out)[2][2], uint16_t* rows,
                                          int off, uint32_t base, uint32_t n, int lane) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                bool on = (mask >> (nb * 16 + mb * 4 + r)) & 1u;
                out[mb >> 1][nb][((mb & 1) << 2) + r] = on ? f2bf(acc[nb][mb][r]) : (uint16_t)0;
            }
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        uint32_t s = base + nb * 16 + c;
        if (s >= n) continue;
        uint16_t* r = rows + (size_t)s * kRowLen + off;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const u16x8& p = out[mb >> 1][nb];
            int q = (mb & 1) << 2;
            *reinterpret_cast<uint2*>(r + mb * 16 + 4 * g) =
                make_uint2((uint32_t)p[q] | ((uint32_t)p[q + 1] << 16), (uint32_t)p[q + 2] | ((uint32_t)p[q + 3] << 16));
        }
    }
}

__global__ __launch_bounds__(kMlpBlock, 2) void k_mlp_train(MlpView v, const TrainSample* __restrict__ smp, uint32_t n,
                                                         float inv_b, uint16_t* rows, float* loss, float* ggrad,
                                                         int nan_inject) {
    __shared__ __align__(16) uint16_t sw[kMirror];
    __shared__ float sb[kBiasCount];
    __shared__ float stage[4 * 32 * kStagePitch];
    load_net(v, sw, sb);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, c = lane & 15, g = lane >> 4;
    float* st = stage + wv * 32 * kStagePitch;
    const uint32_t n_tiles = (n + 127u) / 128u;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint32_t base = tile * 128u + (uint32_t)wv * 32u;
        u16x8 x[3][2];
        float p01[2][3];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            uint32_t s = base + nb * 16 + c;
            bool valid = s < n;
            TrainSample ts{};
            if (valid) ts = smp[s];
            scene_unit(v, F3{ts.x_w.x, ts.x_w.y, ts.x_w.z}, p01[nb]);
            u16x8 e[3];
            encode_slots(v, valid, p01[nb], F3{ts.n_pb.x, ts.n_pb.y, ts.n_pb.z}, F3{ts.wo.x, ts.wo.y, ts.wo.z}, g, e);
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) {
                x[kb][nb] = e[kb];
                if (valid) *reinterpret_cast<u16x8*>(rows + (size_t)s * kRowLen + kRowX + kb * 32 + g * 8) = e[kb];
            }
        }
        FwdOut f;
        tile_forward(sw, sb, x, f, rows, base, n, lane);
        stage_outputs(f, sb, st, lane);
        __syncthreads();
        if (lane < 32) {
            uint32_t s = base + lane;
            float* o = st + lane * kStagePitch;
            if (s < n) {
                TrainSample ts = smp[s];
                float lv = head_loss_grad(o, F3{ts.wi.x, ts.wi.y, ts.wi.z}, ts.n_pb.w, ts.x_w.w, inv_b);
                if (nan_inject && s == 0) o[0] = __int_as_float(0x7fc00000);
                loss[s] = lv;
                uint16_t* r = rows + (size_t)s * kRowLen + kRowDO;
                for (int k = 0; k < kOutputs; k += 2)
                    *reinterpret_cast<uint32_t*>(r + k) = (uint32_t)f2bf(o[k]) | ((uint32_t)f2bf(o[k + 1]) << 16);
            } else {
                for (int k = 0; k < kOutputs; ++k) o[k] = 0.f;
            }
        }
        __syncthreads();
        u16x8 d[2][2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    int o = log_of(kb * 32 + 8 * g + j);
                    d[kb][nb][j] = o < kOutputs ? f2bf(st[(nb * 16 + c) * kStagePitch + o]) : (uint16_t)0;
                }
        f32x4 acc[2][4];
        layer_bwd<4>(sw + kMW4, 64, kOutputs, false, d, acc, lane);
        mask_pack(acc, f.m3, d, rows, kRowD3, base, n, lane);
        layer_bwd<4>(sw + kMW3, 64, 64, false, d, acc, lane);
        mask_pack(acc, f.m2, d, rows, kRowD2, base, n, lane);
        layer_bwd<4>(sw + kMW2, 64, 64, false, d, acc, lane);
        mask_pack(acc, f.m1, d, rows, kRowD1, base, n, lane);
        f32x4 ax[2][2];
        layer_bwd<2>(sw + kMW1, 96, 64, true, d, ax, lane);
        // hash-grid gradients: this lane holds inputs 8g + 4mb + r = levels 4g + 2mb + r/2.
        // The 16 lanes of a row hold 16 consecutive samples at one level; each run of
        // lanes that hit the same entry adds its sum with one atomic pair from its last
        // lane (spatially sorted batches make such runs common on the coarse levels).
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const bool valid = base + nb * 16 + c < n;
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int lp = 0; lp < 2; ++lp) {
                    const int l = 4 * g + 2 * mb + lp;
                    const float g0 = ax[nb][mb][2 * lp], g1 = ax[nb][mb][2 * lp + 1];
                    const bool live = valid && (g0 != 0.f || g1 != 0.f);
                    uint32_t ic[3];
                    float fr[3];
                    grid_cell(v, l, p01[nb], ic, fr);
                    float* tab = ggrad + v.goff[l] + (size_t)(blockIdx.x % v.gcopies[l]) * v.gents[l] * 2u;
#pragma unroll
                    for (int cc = 0; cc < 8; ++cc) {
                        const uint32_t e = grid_entry(v, l, ic[0] + (cc & 1), ic[1] + ((cc >> 1) & 1), ic[2] + (cc >> 2));
                        const float w = corner_weight(cc, fr);
                        const bool on = live && w != 0.f;
                        float a = on ? w * g0 : 0.f, b = on ? w * g1 : 0.f;
                        if (row_run_sum(on ? e : 0xFFFFFFFFu, c, a, b) && on) {
                            atomicAdd(tab + 2 * e, a);
                            atomicAdd(tab + 2 * e + 1, b);
                        }
                    }
                }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ weight gradients
constexpr int kTP = 40;  // LDS pitch (BF16) of one feature row: 32 samples + 8 pad

struct WTile {
    int drow, arow, off, ld;
};
__device__ __forceinline__ WTile wtile(int k) {
    // 64 tiles of 16x16: W1 4x5 (inputs 0..79), W2 4x4, W3 4x4, W4 3x4
    if (k < 20) return WTile{kRowD1 + (k / 5) * 16, kRowX + (k % 5) * 16, kOffW1 + (k / 5) * 16 * 96 + (k % 5) * 16, 96};
    if (k < 36) { k -= 20; return WTile{kRowD2 + (k / 4) * 16, kRowH1 + (k % 4) * 16, kOffW2 + (k / 4) * 16 * 64 + (k % 4) * 16, 64}; }
    if (k < 52) { k -= 36; return WTile{kRowD3 + (k / 4) * 16, kRowH2 + (k % 4) * 16, kOffW3 + (k / 4) * 16 * 64 + (k % 4) * 16, 64}; }
    k -= 52;
    return WTile{kRowDO + (k / 4) * 16, kRowH3 + (k % 4) * 16, kOffW4 + (k / 4) * 16 * 64 + (k % 4) * 16, 64};
}

__global__ __launch_bounds__(256) void k_mlp_wgrad(const uint16_t* __restrict__ rows, uint32_t n, uint32_t per_group,
                                                   float* partial, int accumulate) {
    __shared__ __align__(16) uint16_t T[kRowLen * kTP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, c = lane & 15, g = lane >> 4;
    const uint32_t s0 = blockIdx.x * per_group;
    const uint32_t s1 = min(n, s0 + per_group);
    f32x4 acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int bfeat = tid < 64 ? kRowD1 + tid : tid < 128 ? kRowD2 + (tid - 64) : tid < 192 ? kRowD3 + (tid - 128)
                                                                                   : kRowDO + (tid - 192);
    float bacc = 0.f;
    for (uint32_t cs = s0; cs < s1; cs += 32) {
        __syncthreads();
        for (int u = tid; u < 32 * (kRowLen / 8); u += 256) {
            int sl = u / (kRowLen / 8), part = u % (kRowLen / 8);
            uint32_t s = cs + (uint32_t)sl;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (s < s1) val = reinterpret_cast<const uint4*>(rows + (size_t)s * kRowLen)[part];
            const uint32_t w4[4] = {val.x, val.y, val.z, val.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                T[(part * 8 + 2 * q) * kTP + sl] = (uint16_t)(w4[q] & 0xffffu);
                T[(part * 8 + 2 * q + 1) * kTP + sl] = (uint16_t)(w4[q] >> 16);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            WTile t = wtile(wv * 16 + k);
            u16x8 a = *reinterpret_cast<const u16x8*>(T + (t.drow + c) * kTP + 8 * g);
            u16x8 b = *reinterpret_cast<const u16x8*>(T + (t.arow + c) * kTP + 8 * g);
            acc[k] = mfma16(a, b, acc[k]);
        }
        if (tid < kBiasCount)
            for (int s = 0; s < 32; ++s) bacc += bf2f(T[bfeat * kTP + s]);
    }
    float* P = partial + (size_t)blockIdx.x * kParams;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        WTile t = wtile(wv * 16 + k);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int idx = t.off + (4 * g + r) * t.ld + c;
            P[idx] = (accumulate ? P[idx] : 0.f) + acc[k][r];
        }
    }
    if (tid < kBiasCount) {
        int off = tid < 64 ? kOffB1 + tid : tid < 128 ? kOffB2 + (tid - 64) : tid < 192 ? kOffB3 + (tid - 128)
                                                                                : kOffB4 + (tid - 192);
        P[off] = (accumulate ? P[off] : 0.f) + bacc;
    }
    if (!accumulate)
        for (int i = tid; i < kHidden * (kInputs - kInputsUsed); i += 256)
            P[kOffW1 + (i / 16) * kInputs + kInputsUsed + (i % 16)] = 0.f;
}

// ------------------------------------------------------------------ optimiser
__device__ __forceinline__ int mirror_index(int p) {
    if (p < kOffB1) return kMW1 + p;
    if (p >= kOffW2 && p < kOffB2) { int q = p - kOffW2; return kMW2 + (q / 64) * 64 + phys_of(q % 64); }
    if (p >= kOffW3 && p < kOffB3) { int q = p - kOffW3; return kMW3 + (q / 64) * 64 + phys_of(q % 64); }
    if (p >= kOffW4 && p < kOffB4) { int q = p - kOffW4; return kMW4 + (q / 64) * 64 + phys_of(q % 64); }
    return -1;
}

__global__ void k_mlp_adam(float* params, float* m, float* v, uint16_t* mirror, const float* __restrict__ partial,
                           uint32_t groups, AdamParams ap, int update, float* grad_out, uint32_t* status) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float gs = 0.f;
    for (uint32_t w = 0; w < groups; ++w) gs += partial[(size_t)w * kParams + p];
    if (grad_out) grad_out[p] = gs;
    if (!isfinite(gs)) {
        atomicOr(status, 1u);
        return;
    }
    if (!update) return;
    float mm = ap.beta1 * m[p] + (1.f - ap.beta1) * gs;
    float vv = ap.beta2 * v[p] + (1.f - ap.beta2) * (gs * gs);
    m[p] = mm;
    v[p] = vv;
    float th = params[p] - ap.lr * (mm / ap.bc1) / (sqrtf(vv / ap.bc2) + ap.eps);
    params[p] = th;
    int mi = mirror_index(p);
    if (mi >= 0) mirror[mi] = f2bf(th);
}

// Sums (in copy order) and clears the private gradient copies of each hash entry,
// then applies Adam; parameter p = (level, entry, feature) in the [level][table][2] layout.
__global__ void k_grid_adam(MlpView view, float* master, uint16_t* bf, float* m, float* v, float* grad, size_t n,
                            AdamParams ap, int update, float* grad_out, uint32_t* status) {
    const size_t per_level = (size_t)2 << view.log2_table;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(p / per_level);
        const size_t q = p - (size_t)l * per_level;
        float gs = 0.f;
        if (q < (size_t)view.gents[l] * 2u) {
            float* gp = grad + view.goff[l] + q;
            const size_t cstride = (size_t)view.gents[l] * 2u;
            for (uint32_t c = 0; c < view.gcopies[l]; ++c) {
                gs += gp[c * cstride];
                gp[c * cstride] = 0.f;
            }
        }
        if (grad_out) grad_out[p] = gs;
        if (!isfinite(gs)) {
            atomicOr(status, 2u);
            continue;
        }
        if (!update) continue;
        float mm = m[p], vv = v[p];
        if (gs == 0.f && mm == 0.f && vv == 0.f) continue;
        mm = ap.beta1 * mm + (1.f - ap.beta1) * gs;
        vv = ap.beta2 * vv + (1.f - ap.beta2) * (gs * gs);
        m[p] = mm;
        v[p] = vv;
        float th = master[p] - ap.lr * (mm / ap.bc1) / (sqrtf(vv / ap.bc2) + ap.eps);
        master[p] = th;
        bf[p] = f2bf(th);
    }
}

__global__ void k_mlp_mirror(const float* params, uint16_t* mirror) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    int mi = mirror_index(p);
    if (mi >= 0) mirror[mi] = f2bf(params[p]);
}

__global__ void k_grid_mirror(const float* master, uint16_t* bf, size_t n) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
        bf[p] = f2bf(master[p]);
}

// Deterministic sum of in[i * stride]: each of kSumBlocks blocks reduces a fixed
// interleaved share in FP64 into part[block]; k_sum_final adds the parts in order.
constexpr int kSumBlocks = 256;
__global__ __launch_bounds__(256) void k_sum_part(const float* in, size_t n, size_t stride, double* part) {
    __shared__ double red[256];
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kSumBlocks * 256)
        a += (double)in[i * stride];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_sum_final(const double* part, float* out) {
    __shared__ double red[256];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)red[0];
}

// ------------------------------------------------------------------ samples and inference
__global__ void k_mlp_gather(const Record* __restrict__ rec, const TrainAux* __restrict__ aux, uint32_t n_rec,
                             uint32_t offset, uint32_t stride, uint32_t batch, float w_scale, float w_clip,
                             TrainSample* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= batch) return;
    uint32_t r = (uint32_t)(((uint64_t)i * stride + offset) % n_rec);
    const float* v = rec[r].v;
    TrainAux a = aux[r];
    float w = a.a.w * w_scale;
    w = (isfinite(w) && w > 0.f) ? fminf(w, w_clip) : 0.f;
    TrainSample t;
    t.x_w = make_float4(v[0], v[1], v[2], w);
    t.n_pb = make_float4(v[3], v[4], v[5], a.b.x);
    t.wo = make_float4(a.a.x, a.a.y, a.a.z, 0.f);
    t.wi = make_float4(v[6], v[7], v[8], 0.f);
    out[i] = t;
}

// Spatial order of a training batch: a 24-bit Morton key of each sample's position in
// the scene cube (8 bits per axis), radix-sorted, then the samples permuted into it.
__device__ __forceinline__ uint32_t spread_bits8(uint32_t x) {
    x &= 0xffu;
    x = (x | (x << 8)) & 0x0000f00fu;
    x = (x | (x << 4)) & 0x000c30c3u;
    x = (x | (x << 2)) & 0x00249249u;
    return x;
}

__global__ void k_mlp_sample_keys(MlpView v, const TrainSample* __restrict__ smp, uint32_t n, uint32_t* keys,
                                  uint32_t* vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = smp[i].x_w;
    float p01[3];
    scene_unit(v, F3{x.x, x.y, x.z}, p01);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = min((uint32_t)(p01[a] * 256.f), 255u);
    keys[i] = (spread_bits8(q[0]) << 2) | (spread_bits8(q[1]) << 1) | spread_bits8(q[2]);
    vals[i] = i;
}

__global__ void k_mlp_permute(const TrainSample* __restrict__ in, const uint32_t* __restrict__ order, uint32_t n,
                              TrainSample* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = in[order[i]];
}

// Forward + mixture for n explicit points.
__global__ __launch_bounds__(kMlpBlock) void k_mlp_eval(MlpView v, const float4* __restrict__ xs,
                                                        const float4* __restrict__ ns, const float4* __restrict__ wos,
                                                        uint32_t n, float4* mix) {
    __shared__ __align__(16) uint16_t sw[kMirror];
    __shared__ float sb[kBiasCount];
    __shared__ float stage[4 * 32 * kStagePitch];
    load_net(v, sw, sb);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, c = lane & 15, g = lane >> 4;
    float* st = stage + wv * 32 * kStagePitch;
    const uint32_t n_tiles = (n + 127u) / 128u;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint32_t base = tile * 128u + (uint32_t)wv * 32u;
        u16x8 x[3][2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            uint32_t s = base + nb * 16 + c;
            bool valid = s < n;
            float4 p4 = valid ? xs[s] : make_float4(0.f, 0.f, 0.f, 0.f);
            float4 n4 = valid ? ns[s] : make_float4(0.f, 0.f, 1.f, 0.f);
            float4 w4 = valid ? wos[s] : make_float4(0.f, 0.f, 1.f, 0.f);
            float p01[3];
            scene_unit(v, F3{p4.x, p4.y, p4.z}, p01);
            u16x8 e[3];
            encode_slots(v, valid, p01, F3{n4.x, n4.y, n4.z}, F3{w4.x, w4.y, w4.z}, g, e);
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) x[kb][nb] = e[kb];
        }
        FwdOut f;
        tile_forward(sw, sb, x, f, nullptr, base, n, lane);
        stage_outputs(f, sb, st, lane);
        __syncthreads();
        if (lane < 32 && base + lane < n) head_mixture(st + lane * kStagePitch, mix + (size_t)(base + lane) * kMixF4);
        __syncthreads();
    }
}

// Mixture for every shading point of a bounce, by queue slot (c = 0 where the
// point is not guided: misses, lights, specular materials). With MULTI (batches of
// views over several scenes) only the slots whose path belongs to scene `scene` are
// written, with that scene's network; the other slots are left as they are.
template <bool MULTI>
__global__ __launch_bounds__(kMlpBlock) void k_mlp_infer(DevScene s0, PathState ps, MlpView v, int parity, float4* mix,
                                                         uint32_t scene, const uint32_t* __restrict__ order) {
    __shared__ __align__(16) uint16_t sw[kMirror];
    __shared__ float sb[kBiasCount];
    __shared__ float stage[4 * 16 * kStagePitch];  // half a tile at a time: 3 blocks per CU
    const uint32_t n = ps.counters[parity];
    const uint32_t n_tiles = (n + 127u) / 128u;
    uint32_t t_begin = blockIdx.x, t_end = n_tiles, t_step = gridDim.x;
    if (order && (gridDim.x & 7u) == 0u) {
        // Workgroups are dispatched to the 8 XCDs round robin (block b runs on XCD
        // b % 8): the blocks of XCD x take the x-th eighth of the Morton-ordered
        // queries, so the hash-grid lines of one eighth of the scene's surface stay in
        // that XCD's L2 instead of every XCD caching lines of the whole scene.
        const uint32_t xcd = blockIdx.x & 7u;
        t_begin = (uint32_t)(((uint64_t)n_tiles * xcd) / 8u) + (blockIdx.x >> 3);
        t_end = (uint32_t)(((uint64_t)n_tiles * (xcd + 1u)) / 8u);
        t_step = gridDim.x >> 3;
    }
    if (t_begin >= t_end) return;
    load_net(v, sw, sb);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, c = lane & 15, g = lane >> 4;
    float* st = stage + wv * 16 * kStagePitch;
    for (uint32_t tile = t_begin; tile < t_end; tile += t_step) {
        const uint32_t base = tile * 128u + (uint32_t)wv * 32u;
        u16x8 x[3][2];
        bool valid_nb[2], mine_nb[2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const uint32_t k = base + nb * 16 + c;
            bool valid = k < n, mine = true;
            const uint32_t i = (valid && order) ? order[k] : k;
            F3 xp{0.f, 0.f, 0.f}, sn{0.f, 0.f, 1.f}, wo{0.f, 0.f, 1.f};
            const DevScene* sp = &s0;
            if (MULTI && valid) {
                const uint32_t path = ps.queue[parity][i];
                const uint32_t sid = ps.bviews[ps.path_view[path]].scene;
                mine = sid == scene;
                sp = &ps.bscenes[sid];
                valid = mine;
            }
            const DevScene& s = *sp;
            if (valid) {
                float4 h = ps.hit[i];
                uint32_t hid = __float_as_uint(h.x);
                valid = hid != kMiss;
                if (valid) {
                    const TriShade& tsh = s.shade[hid];
                    const MaterialRec& m = s.mats[tsh.mat];
                    valid = !m.specular && !m.is_light;
                    if (valid) {
                        float4 ro = ps.ray_o[parity][i], rdv = ps.ray_d[parity][i];
                        F3 rd{rdv.x, rdv.y, rdv.z};
                        xp = F3{ro.x, ro.y, ro.z} + rd * h.y;
                        float ha = 1.f - (h.z + h.w);
                        sn = normalize((ld3(tsh.n0) * ha + ld3(tsh.n1) * h.z) + ld3(tsh.n2) * h.w);
                        wo = -rd;
                        if (m.two_sided && dot(wo, sn) < 0.f) sn = -sn;
                    }
                }
            }
            valid_nb[nb] = valid;
            mine_nb[nb] = mine;
            float p01[3];
            scene_unit(v, xp, p01);
            u16x8 e[3];
            encode_slots(v, valid, p01, sn, wo, g, e);
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) x[kb][nb] = e[kb];
        }
        FwdOut f;
        tile_forward(sw, sb, x, f, nullptr, base, n, lane);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            stage_half(f, sb, st, lane, nb);
            // lanes 0..15 hold the validity of samples nb*16 + c: share it through the
            // staging pad (1 guided, 0 unguided point, 2 another scene's slot)
            if (g == 0) st[c * kStagePitch + 48] = !mine_nb[nb] ? 2.f : valid_nb[nb] ? 1.f : 0.f;
            __syncthreads();
            const uint32_t k = base + (uint32_t)nb * 16u + (uint32_t)lane;
            if (lane < 16 && k < n) {
                const uint32_t i = order ? order[k] : k;
                float4* mx = mix + (size_t)i * kMixF4;
                const float flag = st[lane * kStagePitch + 48];
                if (flag == 1.f) head_mixture(st + lane * kStagePitch, mx);
                else if (flag == 0.f) mx[12] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ batches
// Queue slots of a bounce grouped by scene, each scene's group padded to whole
// 128-slot tiles, so that one launch evaluates every scene's network (each tile
// belongs to one scene; a block loads a network when its tiles change scene).
// Wave-aggregated append: the lanes that hold the same key take consecutive entries
// of counter[key] with one atomic per distinct key of the wave (queues hold runs of
// one scene, so most waves issue one atomic). Returns the lane's entry (active lanes).
__device__ __forceinline__ uint32_t wave_key_append(uint32_t* counter, uint32_t key, bool active) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long todo = __ballot(active);
    uint32_t mine = 0;
    while (todo) {
        const int leader = __ffsll(todo) - 1;
        const uint32_t k = (uint32_t)__shfl((int)key, leader);
        const unsigned long long grp = __ballot(active && key == k) & todo;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counter[k], (uint32_t)__popcll(grp));
        base = (uint32_t)__shfl((int)base, leader);
        if (active && key == k && ((grp >> lane) & 1ull)) mine = base + (uint32_t)__popcll(grp & below);
        todo &= ~grp;
    }
    return mine;
}

// The scene of queue slot i, and whether that scene has a network.
__device__ __forceinline__ bool slot_scene(const PathState& ps, int parity, const MlpView* guides, uint32_t i,
                                           uint32_t n, uint32_t& sid) {
    sid = 0;
    if (i >= n) return false;
    sid = ps.bviews[ps.path_view[ps.queue[parity][i]]].scene;
    return guides[sid].wmirror != nullptr;
}

// Scenes below kLdsScenes are counted per block in LDS and reach the global counters
// with one atomic per block and scene: a chunk's queue is mostly one scene, and one
// global atomic per wave on that one counter serialised the whole launch (3.9 ms for
// 22M slots). Scenes from kLdsScenes up take the per-wave global path.
constexpr uint32_t kLdsScenes = 256;

__global__ __launch_bounds__(256) void k_scene_count(PathState ps, int parity, const MlpView* guides,
                                                     uint32_t n_scenes, uint32_t* cnt) {
    __shared__ uint32_t lcnt[kLdsScenes];
    for (uint32_t t = threadIdx.x; t < kLdsScenes; t += blockDim.x) lcnt[t] = 0u;
    __syncthreads();
    const uint32_t n = ps.counters[parity];
    const uint32_t stride = gridDim.x * blockDim.x;
    // whole waves iterate together (wave_key_append needs every lane)
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        uint32_t sid;
        const bool on = slot_scene(ps, parity, guides, i0 + threadIdx.x, n, sid);
        (void)wave_key_append(lcnt, sid, on && sid < kLdsScenes);
        (void)wave_key_append(cnt, sid, on && sid >= kLdsScenes);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kLdsScenes && t < n_scenes; t += blockDim.x)
        if (lcnt[t]) atomicAdd(&cnt[t], lcnt[t]);
}
// One thread: padded offsets, scatter cursors, tile count; clears the counts.
__global__ void k_scene_offsets(uint32_t n_scenes, uint32_t* cnt, uint32_t* off, uint32_t* cur, uint32_t* n_tiles) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t o = 0;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        off[s] = o;
        cur[s] = o;
        o += (cnt[s] + 127u) & ~127u;
        cnt[s] = 0u;
    }
    off[n_scenes] = o;
    *n_tiles = o / 128u;
}
// Padding entries of the grouped order stay ~0 (no slot).
__global__ void k_perm_clear(const uint32_t* n_tiles, uint32_t* perm) {
    const uint32_t total = *n_tiles * 128u;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) perm[j] = ~0u;
}
// Each block owns a contiguous run of whole waves of the queue: it counts its slots
// per scene in LDS, reserves each scene's entries with one atomic on that scene's
// cursor, then hands them out from LDS cursors in a second sweep. Slots of scenes
// without a network get c = 0 (unguided) here; k_mlp_infer_grouped writes every other
// slot of the queue.
__global__ __launch_bounds__(256) void k_scene_scatter(PathState ps, int parity, const MlpView* guides,
                                                       uint32_t n_scenes, uint32_t* cur, uint32_t* perm,
                                                       float4* mix) {
    __shared__ uint32_t lcur[kLdsScenes];
    for (uint32_t t = threadIdx.x; t < kLdsScenes; t += blockDim.x) lcur[t] = 0u;
    __syncthreads();
    const uint32_t n = ps.counters[parity];
    const uint32_t per = ((n + gridDim.x - 1u) / gridDim.x + 63u) & ~63u;
    const uint32_t lo = min(blockIdx.x * per, n), hi = min(lo + per, n);
    for (uint32_t i0 = lo; i0 < hi; i0 += blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t sid;
        const bool on = slot_scene(ps, parity, guides, i, hi, sid);
        if (!on && i < hi) mix[(size_t)i * kMixF4 + 12] = make_float4(0.f, 0.f, 0.f, 0.f);
        (void)wave_key_append(lcur, sid, on && sid < kLdsScenes);
        const bool wide = on && sid >= kLdsScenes;
        const uint32_t at = wave_key_append(cur, sid, wide);
        if (wide) perm[at] = i;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kLdsScenes && t < n_scenes; t += blockDim.x)
        if (lcur[t]) lcur[t] = atomicAdd(&cur[t], lcur[t]);
    __syncthreads();
    for (uint32_t i0 = lo; i0 < hi; i0 += blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t sid;
        const bool on = slot_scene(ps, parity, guides, i, hi, sid) && sid < kLdsScenes;
        const uint32_t at = wave_key_append(lcur, sid, on);
        if (on) perm[at] = i;
    }
}

// Every scene's network over its own slots in one launch (see k_scene_count): blocks
// stride over the tiles, so the blocks in flight work on neighbouring tiles (slots of
// neighbouring pixels, whose hash-grid lines the L2 then shares); a block's tiles only
// move forward, so it loads each scene's network at most once. The slots of scenes
// without a network were set unguided by k_scene_scatter.
__global__ __launch_bounds__(kMlpBlock) void k_mlp_infer_grouped(PathState ps, int parity,
                                                                 const MlpView* __restrict__ guides, uint32_t n_scenes,
                                                                 const uint32_t* __restrict__ off,
                                                                 const uint32_t* __restrict__ n_tiles_p,
                                                                 const uint32_t* __restrict__ perm, float4* mix) {
    __shared__ __align__(16) uint16_t sw[kMirror];
    __shared__ float sb[kBiasCount];
    __shared__ float stage[4 * 16 * kStagePitch];
    __shared__ uint32_t sslot[4 * 16];
    const uint32_t n_tiles = *n_tiles_p;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, c = lane & 15, g = lane >> 4;
    float* st = stage + wv * 16 * kStagePitch;
    int loaded = -1;
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        // the tile's scene: the last s with off[s] <= tile * 128
        uint32_t lo = 0, hi = n_scenes;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= tile * 128u) lo = mid;
            else hi = mid;
        }
        // block-uniform: the view's fields are read with scalar loads, as k_mlp_infer
        // reads its view from the kernel arguments (a copy of the view in LDS made every
        // hash-grid address a per-lane LDS read: 79.5 -> 68.6 ms/SPP for the 48-view
        // batch guided at every bounce)
        const int scene = __builtin_amdgcn_readfirstlane((int)lo);
        const MlpView& sv = guides[scene];
        if (scene != loaded) {
            __syncthreads();  // every wave is done with the previous network
            load_net(sv, sw, sb);
            loaded = scene;
        }
        const DevScene& s = ps.bscenes[scene];
        const uint32_t base = tile * 128u + (uint32_t)wv * 32u;
        u16x8 x[3][2];
        bool valid_nb[2];
        uint32_t slot_nb[2];
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const uint32_t i = perm[base + nb * 16 + c];
            slot_nb[nb] = i;
            bool valid = i != ~0u;
            F3 xp{0.f, 0.f, 0.f}, sn{0.f, 0.f, 1.f}, wo{0.f, 0.f, 1.f};
            if (valid) {
                float4 h = ps.hit[i];
                uint32_t hid = __float_as_uint(h.x);
                valid = hid != kMiss;
                if (valid) {
                    const TriShade& tsh = s.shade[hid];
                    const MaterialRec& m = s.mats[tsh.mat];
                    valid = !m.specular && !m.is_light;
                    if (valid) {
                        float4 ro = ps.ray_o[parity][i], rdv = ps.ray_d[parity][i];
                        F3 rd{rdv.x, rdv.y, rdv.z};
                        xp = F3{ro.x, ro.y, ro.z} + rd * h.y;
                        float ha = 1.f - (h.z + h.w);
                        sn = normalize((ld3(tsh.n0) * ha + ld3(tsh.n1) * h.z) + ld3(tsh.n2) * h.w);
                        wo = -rd;
                        if (m.two_sided && dot(wo, sn) < 0.f) sn = -sn;
                    }
                }
            }
            valid_nb[nb] = valid;
            float p01[3];
            scene_unit(sv, xp, p01);
            u16x8 e[3];
            encode_slots(sv, valid, p01, sn, wo, g, e);
#pragma unroll
            for (int kb = 0; kb < 3; ++kb) x[kb][nb] = e[kb];
        }
        FwdOut f;
        tile_forward(sw, sb, x, f, nullptr, base, ~0u, lane);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            stage_half(f, sb, st, lane, nb);
            if (g == 0) {
                st[c * kStagePitch + 48] = valid_nb[nb] ? 1.f : 0.f;
                sslot[wv * 16 + c] = slot_nb[nb];
            }
            __syncthreads();
            if (lane < 16) {
                const uint32_t i = sslot[wv * 16 + lane];
                if (i != ~0u) {
                    float4* mx = mix + (size_t)i * kMixF4;
                    if (st[lane * kStagePitch + 48] == 1.f) head_mixture(st + lane * kStagePitch, mx);
                    else mx[12] = make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
            __syncthreads();
        }
    }
}

__global__ void k_mix_query(const float4* __restrict__ mix, uint32_t n, const float4* __restrict__ wi_eval,
                            const float4* __restrict__ u, float* pdf_eval, float4* wi_out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* mx = mix + (size_t)i * kMixF4;
    if (wi_eval) {
        float4 w = wi_eval[i];
        pdf_eval[i] = mix_pdf(mx, F3{w.x, w.y, w.z});
    }
    if (u) {
        float4 uu = u[i];
        F3 d = mix_sample(mx, uu.x, uu.y, uu.z);
        wi_out[i] = make_float4(d.x, d.y, d.z, mix_pdf(mx, d));
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launch_mlp_gather(const Record* rec, const TrainAux* aux, uint32_t n_rec, uint32_t offset, uint32_t stride,
                             uint32_t batch, float w_scale, float w_clip, TrainSample* out, hipStream_t stream) {
    if (!batch || !n_rec) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_gather, dim3((batch + 255) / 256), dim3(256), 0, stream, rec, aux, n_rec, offset, stride,
                       batch, w_scale, w_clip, out);
    return hipGetLastError();
}

hipError_t launch_mlp_sort_batch(const MlpView& v, const TrainSample* in, uint32_t n, const SortBuffers& sb,
                                 TrainSample* out, hipStream_t stream) {
    if (!n) return hipSuccess;
    const dim3 grid((n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_sample_keys, grid, dim3(256), 0, stream, v, in, n, sb.keys, sb.vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(sb.keys, sb.vals, sb.keys_tmp, sb.vals_tmp, n, 24, sb.hist, sb.scan, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_permute, grid, dim3(256), 0, stream, in, sb.vals, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_train_chunk(const MlpView& v, const TrainSample* smp, uint32_t n, float inv_batch,
                                  uint16_t* rows, float* loss, float* grid_grad, int nan_inject, int cu_count,
                                  hipStream_t stream) {
    if (!n) return hipSuccess;
    uint32_t tiles = (n + 127u) / 128u;
    uint32_t grid = std::min<uint32_t>(tiles, (uint32_t)cu_count * 2u);
    hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(kMlpBlock), 0, stream, v, smp, n, inv_batch, rows, loss,
                       grid_grad, nan_inject);
    return hipGetLastError();
}

uint32_t mlp_wgrad_groups(uint32_t n) {
    uint32_t g = (n + 511u) / 512u;
    return std::max<uint32_t>(1u, std::min<uint32_t>(g, 512u));
}

hipError_t launch_mlp_wgrad(const uint16_t* rows, uint32_t n, uint32_t groups, float* partial, int accumulate,
                            hipStream_t stream) {
    uint32_t per = (n + groups - 1u) / groups;
    per = (per + 31u) & ~31u;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(groups), dim3(256), 0, stream, rows, n, per, partial, accumulate);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* params, float* m, float* v, uint16_t* mirror, const float* partial, uint32_t groups,
                           const AdamParams& ap, int update, float* grad_out, uint32_t* status, hipStream_t stream) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, stream, params, m, v, mirror, partial,
                       groups, ap, update, grad_out, status);
    return hipGetLastError();
}

hipError_t launch_grid_adam(const MlpView& view, float* grid_master, uint32_t* grid_bf, float* m, float* v, float* grad,
                            size_t n, const AdamParams& ap, int update, float* grad_out, uint32_t* status,
                            hipStream_t stream) {
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_grid_adam, dim3(blocks), dim3(256), 0, stream, view, grid_master,
                       reinterpret_cast<uint16_t*>(grid_bf), m, v, grad, n, ap, update, grad_out, status);
    return hipGetLastError();
}

hipError_t launch_sum(const float* in, size_t n, size_t stride, double* part, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_sum_part, dim3(kSumBlocks), dim3(256), 0, stream, in, n, stride, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, stream, part, out);
    return hipGetLastError();
}

hipError_t launch_mlp_mirror(const float* params, uint16_t* mirror, hipStream_t stream) {
    hipLaunchKernelGGL(k_mlp_mirror, dim3((kParams + 255) / 256), dim3(256), 0, stream, params, mirror);
    return hipGetLastError();
}

hipError_t launch_grid_mirror(const float* grid_master, uint32_t* grid_bf, size_t n, hipStream_t stream) {
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_grid_mirror, dim3(blocks), dim3(256), 0, stream, grid_master,
                       reinterpret_cast<uint16_t*>(grid_bf), n);
    return hipGetLastError();
}

hipError_t launch_mlp_eval(const MlpView& v, const float4* x, const float4* nrm, const float4* wo, uint32_t n,
                           float4* mix, int cu_count, hipStream_t stream) {
    if (!n) return hipSuccess;
    uint32_t tiles = (n + 127u) / 128u;
    uint32_t grid = std::min<uint32_t>(tiles, (uint32_t)cu_count * 2u);
    hipLaunchKernelGGL(k_mlp_eval, dim3(grid), dim3(kMlpBlock), 0, stream, v, x, nrm, wo, n, mix);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const DevScene& s, const PathState& ps, const MlpView& v, int parity, float4* mix,
                            uint32_t grid, hipStream_t stream, bool multi, uint32_t scene, const uint32_t* order) {
    if (multi)
        hipLaunchKernelGGL(k_mlp_infer<true>, dim3(grid), dim3(kMlpBlock), 0, stream, s, ps, v, parity, mix, scene,
                           order);
    else
        hipLaunchKernelGGL(k_mlp_infer<false>, dim3(grid), dim3(kMlpBlock), 0, stream, s, ps, v, parity, mix, 0u,
                           order);
    return hipGetLastError();
}

hipError_t launch_mlp_infer_grouped(const PathState& ps, int parity, const MlpView* guides, uint32_t n_scenes,
                                    const GroupBuffers& gb, float4* mix, uint32_t grid, hipStream_t stream) {
    hipLaunchKernelGGL(k_scene_count, dim3(grid), dim3(256), 0, stream, ps, parity, guides, n_scenes, gb.cnt);
    hipLaunchKernelGGL(k_scene_offsets, dim3(1), dim3(64), 0, stream, n_scenes, gb.cnt, gb.off, gb.cur, gb.n_tiles);
    hipLaunchKernelGGL(k_perm_clear, dim3(grid), dim3(256), 0, stream, gb.n_tiles, gb.perm);
    hipLaunchKernelGGL(k_scene_scatter, dim3(grid), dim3(256), 0, stream, ps, parity, guides, n_scenes, gb.cur,
                       gb.perm, mix);
    hipLaunchKernelGGL(k_mlp_infer_grouped, dim3(grid), dim3(kMlpBlock), 0, stream, ps, parity, guides, n_scenes,
                       gb.off, gb.n_tiles, gb.perm, mix);
    return hipGetLastError();
}

void mlp_view_frame(MlpView& v, uint32_t log2_table, float res_min, float res_max, const float* centre, float radius) {
    // levels res_l = floor(res_min b^l), b = (res_max / res_min)^(1/15); dense indexing
    // while (res_l + 1)^3 fits the table
    const size_t T = (size_t)1 << log2_table;
    v.table_mask = (uint32_t)(T - 1);
    v.log2_table = log2_table;
    const double b = std::exp((std::log((double)res_max) - std::log((double)res_min)) / (kHashLevels - 1));
    for (int l = 0; l < kHashLevels; ++l) {
        double r = std::floor((double)res_min * std::pow(b, (double)l) + 1e-6);
        v.res[l] = (uint32_t)std::max(1.0, r);
        double cells = std::pow((double)v.res[l] + 1.0, 3.0);
        v.dense[l] = cells <= (double)T ? 1u : 0u;
    }
    for (int a = 0; a < 3; ++a) v.lo[a] = centre[a] - radius;
    v.inv_ext = radius > 0.f ? 1.f / (2.f * radius) : 0.f;
}

hipError_t launch_mix_query(const float4* mix, uint32_t n, const float4* wi_eval, const float4* u, float* pdf_eval,
                            float4* wi_out, hipStream_t stream) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mix_query, dim3((n + 255) / 256), dim3(256), 0, stream, mix, n, wi_eval, u, pdf_eval, wi_out);
    return hipGetLastError();
}

}  // namespace pg
