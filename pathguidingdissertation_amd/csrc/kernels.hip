// kernels.hip — gfx950 kernels of the wavefront path tracer.
//
// Pipeline per sample pass (one path per pixel): raygen -> for each bounce
// { extend (closest hit) -> shade (emission/MIS, NEE, Russian roulette, BSDF
// sampling; appends to the shadow and next-ray queues) -> shadow (any hit) } ->
// resolve (box-filter splat into the film).
//
// Traversal follows BVHNode::traverse / traverseVisible (Geometry.h:448-567)
// operation for operation: same slab arithmetic with std::min/std::max NaN
// semantics, same near-first order with ties to the left child, same pruning,
// same strict `t < best` update, so triangle IDs and barycentrics are bit-exact
// against the reference on the same BVH. Per-lane traversal stacks live in LDS
// (stride = block size, bank-conflict free); 32-byte nodes are read as child
// pairs (64 contiguous bytes); triangles are stored in leaf order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_math.h"
#include "kernels.h"

#include "device_shading.h"
#include "guide_device.h"
#include "guiding.h"
#include "mlp_device.h"

namespace pg {

// The traversal stack of this lane in a render kernel: STACK entries in the block's
// LDS, deeper ones in ps.spill (one column per lane of the grid).
template <int STACK>
__device__ __forceinline__ SplitStack<STACK> lane_stack(int* lds, const PathState& ps) {
    const uint32_t lane = blockIdx.x * kBlock + threadIdx.x;
    return SplitStack<STACK>{(lds_int*)(lds + threadIdx.x), ps.spill ? (global_int*)(ps.spill + lane) : nullptr,
                             ps.spill_lanes};
}

// ------------------------------------------------------------------ test-API kernels
template <bool COUNT, int STACK>
__global__ __launch_bounds__(kBlock) void k_trace_closest(DevScene s, const float* __restrict__ rays, uint32_t n,
                                                          uint32_t* id, float* t, float* a, float* b, float* g,
                                                          unsigned long long* stats) {
    __shared__ int lds[STACK * kBlock];
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t cn = 0, ct = 0;
    if (i < n) {
        const float* r = rays + 6 * (size_t)i;
        F3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
        uint32_t hid = kMiss;
        float ht = kFltMax, hu = 0.f, hv = 0.f;
        traverse_closest<COUNT>(s, o, d, hid, ht, hu, hv, LdsStack{lds + threadIdx.x}, cn, ct);
        id[i] = hid;
        t[i] = ht;
        a[i] = (hid == kMiss) ? 0.f : 1.f - (hu + hv);
        b[i] = hu;
        g[i] = hv;
    }
    if (COUNT) {
        wave_add_u64(stats + 2, cn);
        wave_add_u64(stats + 3, ct);
    }
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_trace_shadow(DevScene s, const float* __restrict__ rays,
                                                         const float* __restrict__ max_t, uint32_t n, uint8_t* vis) {
    __shared__ int lds[STACK * kBlock];
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 6 * (size_t)i;
    F3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
    uint32_t cn = 0, ct = 0;
    vis[i] = traverse_visible<false>(s, o, d, max_t[i], LdsStack{lds + threadIdx.x}, cn, ct) ? 1 : 0;
}

__global__ void k_generate_rays(DevScene s, const float* pxy, uint32_t n, float* rays) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F3 o, d;
    generate_ray(s.cam, pxy[2 * i], pxy[2 * i + 1], o, d);
    float* r = rays + 6 * (size_t)i;
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = d.x; r[4] = d.y; r[5] = d.z;
}

__global__ void k_rng(uint32_t seed, uint32_t sample, const uint32_t* pixel, const uint32_t* dim, uint32_t n,
                      float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = keyed_rand(path_key(seed, sample, pixel[i]), dim[i]);
}

// ------------------------------------------------------------------ wavefront kernels
// Scene of a work item: the kernel's scene, or in a batch the scene of the path's view.
template <bool MULTI>
__device__ __forceinline__ const DevScene& scene_of(const DevScene& s0, const PathState& ps, uint32_t path) {
    if constexpr (MULTI) return ps.bscenes[ps.bviews[ps.path_view[path]].scene];
    else return s0;
}

__global__ void k_raygen(DevScene s, PathState ps, PassParams p) {
    const uint32_t n = p.width * p.height;
    const uint32_t ns = p.n_samples ? p.n_samples : 1u;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ps.counters[0] = n * ns;
        ps.counters[1] = 0;
        ps.counters[2] = 0;
        ps.counters[3] = 0;  // ray claims of the persistent-lane extend
    }
    if (i >= n * ns) return;
    const uint32_t pix = i % n, smp = i / n;
    uint32_t x = pix % p.width, y = pix / p.width;
    uint32_t key = path_key(p.seed, p.sample + smp, pix);
    float px = (float)x + keyed_rand(key, 0);  // Renderer.h:679-680
    float py = (float)y + keyed_rand(key, 1);
    F3 o, d;
    generate_ray(s.cam, px, py, o, d);
    ps.ray_o[0][i] = make_float4(o.x, o.y, o.z, 0.f);
    ps.ray_d[0][i] = make_float4(d.x, d.y, d.z, 0.f);
    ps.queue[0][i] = i;
    ps.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    ps.radiance[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    ps.film_pos[i] = make_float2(px, py);
    ps.info[i] = make_uint4(key, 2u, 0u, i);
}

template <bool COUNT, int STACK, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_extend(DevScene s0, PathState ps, int parity,
                                                   const uint32_t* __restrict__ order) {
    __shared__ int lds[STACK * kBlock];
    uint32_t n = ps.counters[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.counters[parity ^ 1] = 0;
        ps.counters[2] = 0;
        ps.stats[0] += n;
    }
    const float4* ro = ps.ray_o[parity];
    const float4* rd = ps.ray_d[parity];
    uint32_t cn = 0, ct = 0;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const uint32_t i = order ? order[j] : j;
        float4 a = ro[i], b = rd[i];
        F3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
        uint32_t hid = kMiss;
        float ht = kFltMax, hu = 0.f, hv = 0.f;
        const DevScene& s = scene_of<MULTI>(s0, ps, MULTI ? ps.queue[parity][i] : 0u);
        traverse_closest<COUNT>(s, o, d, hid, ht, hu, hv, lane_stack<STACK>(lds, ps), cn, ct);
        ps.hit[i] = make_float4(__uint_as_float(hid), ht, hu, hv);
    }
    if (COUNT) {
        wave_add_u64(ps.stats + 2, cn);
        wave_add_u64(ps.stats + 3, ct);
    }
}

template <bool COUNT, int STACK, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_shadow(DevScene s0, PathState ps, const uint32_t* __restrict__ order) {
    __shared__ int lds[STACK * kBlock];
    uint32_t n = ps.counters[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.stats[1] += n;
        ps.counters[3] = 0;  // ray claims of the next persistent-lane extend
    }
    uint32_t cn = 0, ct = 0;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const uint32_t i = order ? order[j] : j;
        float4 a = ps.sh_o[i], b = ps.sh_d[i];
        F3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
        const uint32_t path = __float_as_uint(b.w);
        const DevScene& s = scene_of<MULTI>(s0, ps, path);
        if (traverse_visible<COUNT>(s, o, d, a.w, lane_stack<STACK>(lds, ps), cn, ct)) {
            float4 c = ps.sh_c[i];
            float4 L = ps.radiance[path];
            ps.radiance[path] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
        }
    }
    if (COUNT) {
        wave_add_u64(ps.stats + 4, cn);
        wave_add_u64(ps.stats + 5, ct);
    }
}

// Persistent-lane variants of extend and shadow: each wave owns an equal contiguous
// share of the queue, and a lane that finishes its ray claims the next one of its
// wave's share (a wave-uniform cursor, no atomics: device-scope atomics on one counter
// are serialised at memory) whenever fewer than kRefillLanes lanes of the wave are
// busy, so a wave is not held by its slowest ray. Per-ray arithmetic is that of
// traverse_closest / traverse_visible; results are identical.
constexpr int kRefillLanes = 48;
struct WaveShare {
    uint32_t cur, end;
    __device__ __forceinline__ explicit WaveShare(uint32_t n) {
        const uint32_t waves = gridDim.x * (kBlock / 64);
        const uint32_t w = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
        cur = (uint32_t)(((uint64_t)n * w) / waves);
        end = (uint32_t)(((uint64_t)n * (w + 1)) / waves);
    }
    // slot for each wanting lane (in lane order); slots at or past `end` are empty
    __device__ __forceinline__ uint32_t claim(bool want) {
        const unsigned long long m = __ballot(want);
        const int lane = __lane_id();
        const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const uint32_t slot = cur + (uint32_t)__popcll(m & below);
        cur += (uint32_t)__popcll(m);
        return slot;
    }
};

template <bool COUNT, int STACK, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_extend_dyn(DevScene s0, PathState ps, int parity) {
    __shared__ int lds[STACK * kBlock];
    const uint32_t n = ps.counters[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.counters[parity ^ 1] = 0;
        ps.counters[2] = 0;
        ps.stats[0] += n;
    }
    const float4* ro = ps.ray_o[parity];
    const float4* rd = ps.ray_d[parity];
    const SplitStack<STACK> stk = lane_stack<STACK>(lds, ps);
    uint32_t cn = 0, ct = 0;
    TravState t;
    uint32_t slot = 0;
    const DevScene* sc = &s0;
    bool active = false, drained = false;
    WaveShare share(n);
    while (true) {
        if (!drained) {
            const bool want = !active;
            const uint32_t claim = share.claim(want);
            if (want && claim < share.end) {
                slot = claim;
                float4 a = ro[slot], b = rd[slot];
                t.o = F3{a.x, a.y, a.z};
                t.d = F3{b.x, b.y, b.z};
                t.ht = kFltMax;
                t.hid = kMiss;
                t.hu = 0.f;
                t.hv = 0.f;
                if (MULTI) sc = &scene_of<true>(s0, ps, ps.queue[parity][slot]);
                active = trav_begin<COUNT>(*sc, t, cn);
                if (!active) ps.hit[slot] = make_float4(__uint_as_float(kMiss), kFltMax, 0.f, 0.f);
            }
            drained = share.cur >= share.end;
        }
        if (__ballot(active) == 0ull) {
            if (drained) break;
            continue;
        }
        const int thr = drained ? 1 : kRefillLanes;
        do {
            if (active && !trav_step<COUNT>(*sc, t, stk, cn, ct)) {
                ps.hit[slot] = make_float4(__uint_as_float(t.hid), t.ht, t.hu, t.hv);
                active = false;
            }
        } while (__popcll(__ballot(active)) >= thr);
    }
    if (COUNT) {
        wave_add_u64(ps.stats + 2, cn);
        wave_add_u64(ps.stats + 3, ct);
    }
}

template <bool COUNT, int STACK, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_shadow_dyn(DevScene s0, PathState ps) {
    __shared__ int lds[STACK * kBlock];
    const uint32_t n = ps.counters[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps.stats[1] += n;
        ps.counters[3] = 0;  // ray claims of the next persistent-lane extend
    }
    const SplitStack<STACK> stk = lane_stack<STACK>(lds, ps);
    uint32_t cn = 0, ct = 0;
    TravState t;
    uint32_t slot = 0, path = 0;
    const DevScene* sc = &s0;
    bool active = false, drained = false;
    WaveShare share(n);
    while (true) {
        if (!drained) {
            const bool want = !active;
            const uint32_t claim = share.claim(want);
            if (want && claim < share.end) {
                slot = claim;
                float4 a = ps.sh_o[slot], b = ps.sh_d[slot];
                t.o = F3{a.x, a.y, a.z};
                t.d = F3{b.x, b.y, b.z};
                t.ht = a.w;
                path = __float_as_uint(b.w);
                if (MULTI) sc = &scene_of<true>(s0, ps, path);
                active = vis_begin<COUNT>(*sc, t, cn);
                if (!active) {  // misses the root: unoccluded
                    float4 c = ps.sh_c[slot];
                    float4 L = ps.radiance[path];
                    ps.radiance[path] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
                }
            }
            drained = share.cur >= share.end;
        }
        if (__ballot(active) == 0ull) {
            if (drained) break;
            continue;
        }
        const int thr = drained ? 1 : kRefillLanes;
        do {
            if (active) {
                int r = vis_step<COUNT>(*sc, t, stk, cn, ct);
                if (r == 1) {
                    float4 c = ps.sh_c[slot];
                    float4 L = ps.radiance[path];
                    ps.radiance[path] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
                }
                if (r != 0) active = false;
            }
        } while (__popcll(__ballot(active)) >= thr);
    }
    if (COUNT) {
        wave_add_u64(ps.stats + 4, cn);
        wave_add_u64(ps.stats + 5, ct);
    }
}

// Mixture of the quadtrees of the k nearest records' cells at a shading point, each
// neighbour weighted 1/k. `cells` lists the distinct cells as cell | (count << 24),
// terminated by 0xffffffff when fewer than kKnnMax.
__device__ __forceinline__ float guide_pdf_qt(const GuideView& g, const uint32_t* cells, F3 wi) {
    float u, v;
    dir_to_uv(wi.x, wi.y, wi.z, u, v);
    float sum = 0.f;
    uint32_t k = 0;
    for (int j = 0; j < kKnnMax; ++j) {
        uint32_t e = cells[j];
        if (e == 0xffffffffu) break;
        uint32_t cnt = e >> 24;
        sum += (float)cnt * qt_pdf_uv(g.qt + (size_t)(e & 0xffffffu) * g.qt_stride, u, v);
        k += cnt;
    }
    return k ? sum / (float)k * 0.07957747154594767f : 0.f;
}
__device__ __forceinline__ F3 guide_sample_qt(const GuideView& g, const uint32_t* cells, uint32_t key, uint32_t& dim) {
    uint32_t k = 0;
    for (int j = 0; j < kKnnMax && cells[j] != 0xffffffffu; ++j) k += cells[j] >> 24;
    float r0 = keyed_rand(key, dim++), r1 = keyed_rand(key, dim++), r2 = keyed_rand(key, dim++);
    uint32_t t = min((uint32_t)(r0 * (float)k), k - 1u);
    uint32_t cell = cells[0] & 0xffffffu;
    for (int j = 0; j < kKnnMax && cells[j] != 0xffffffffu; ++j) {
        uint32_t cnt = cells[j] >> 24;
        if (t < cnt) { cell = cells[j] & 0xffffffu; break; }
        t -= cnt;
    }
    float u, v, pu;
    qt_sample_uv(g.qt + (size_t)cell * g.qt_stride, r1, r2, u, v, pu);
    F3 w;
    uv_to_dir(u, v, w.x, w.y, w.z);
    return w;
}

// Guide in the BSDF's primary sample space (the design Renderer.h:398-414 outlines:
// nearby records' directions projected into PSS by inverting the BSDF's sampling at
// this vertex, then PSS sampled and mapped back through the BSDF's warp). The k
// nearest records' flux is binned over a 4x4 grid of (u, v); the guide's density in
// PSS is h = 16 w_bin / sum w, and in solid angle p_bsdf(wi) * h(invert(wi)).
struct PssHist {
    float w[16];
    float total;
};
__device__ __forceinline__ bool pss_build(const GuideView& g, const uint32_t* recs, const MaterialRec& m,
                                          const SD& sd, PssHist& h) {
#pragma unroll
    for (int b = 0; b < 16; ++b) h.w[b] = 0.f;
    h.total = 0.f;
    for (int j = 0; j < kKnnMax; ++j) {
        uint32_t r = recs[j];
        if (r == 0xffffffffu) break;
        float4 d = g.rec_dir[r];
        float u, v;
        if (!(d.w > 0.f) || !bsdf_invert(m, sd, F3{d.x, d.y, d.z}, u, v)) continue;
        int b = min((int)(u * 4.f), 3) + 4 * min((int)(v * 4.f), 3);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (q == b) h.w[q] += d.w;
        h.total += d.w;
    }
    return h.total > 0.f;
}
__device__ __forceinline__ float pss_density(const PssHist& h, float u, float v) {
    int b = min((int)(u * 4.f), 3) + 4 * min((int)(v * 4.f), 3);
    float w = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q)
        if (q == b) w = h.w[q];
    return 16.f * w / h.total;
}
__device__ __forceinline__ float guide_pdf_pss(const PssHist* h, const MaterialRec& m, const SD& sd, F3 wi) {
    float u, v;
    if (!bsdf_invert(m, sd, wi, u, v)) return 0.f;
    float pb = 0.f;
    (void)bsdf_eval(m, sd, wi, pb);
    return pb * pss_density(*h, u, v);
}
// A direction from the PSS guide for uniform numbers (u0, r1, r2): u0 picks the bin
// by flux, (r1, r2) place the point inside it, the BSDF's warp maps it back.
__device__ __forceinline__ F3 pss_sample_u(const PssHist* h, const MaterialRec& m, const SD& sd, float u0, float r1,
                                           float r2) {
    float r0 = u0 * h->total;
    int b = 15;
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        float na = acc + h->w[q];
        if (b == 15 && r0 < na && h->w[q] > 0.f) b = q;
        acc = na;
    }
    while (b > 0 && !(h->w[b] > 0.f)) --b;  // r0 rounded past the last bin
    float u = ((float)(b & 3) + r1) * 0.25f, v = ((float)(b >> 2) + r2) * 0.25f;
    return bsdf_warp(m, sd, fminf(u, 0.99999994f), fminf(v, 0.99999994f));
}
__device__ __forceinline__ F3 guide_sample_pss(const PssHist* h, const MaterialRec& m, const SD& sd, uint32_t key,
                                               uint32_t& dim) {
    float u0 = keyed_rand(key, dim++);
    float r1 = keyed_rand(key, dim++), r2 = keyed_rand(key, dim++);
    return pss_sample_u(h, m, sd, u0, r1, r2);
}

// The PSS guide at explicit shading points (material mat, shading normal facing wo,
// the approximate neighbour list of each point): the density of wi_eval and a
// direction sampled with u3 with its density. Points where the guide does not apply
// (specular or non-invertible material, no record flux) get zeros.
__global__ void k_pss_eval(DevScene s, GuideView g, uint32_t mat, const float* __restrict__ pos,
                           const float* __restrict__ nrm, const float* __restrict__ wo,
                           const float* __restrict__ u3, const float* __restrict__ wi_eval, uint32_t n,
                           float* __restrict__ wi_out, float* __restrict__ pdf_sample, float* __restrict__ pdf_eval) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MaterialRec& m = s.mats[mat];
    SD sd;
    sd.x = F3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    sd.sn = normalize(F3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]});
    sd.gn = sd.sn;
    sd.wo = normalize(F3{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]});
    sd.fr = frame_from(sd.sn);
    sd.alb = c3(m.albedo[0], m.albedo[1], m.albedo[2]);
    uint32_t recs[kKnnMax];
    nearby_records(g, sd.x, recs);
    PssHist h;
    float pe = 0.f, ps = 0.f;
    F3 w{0.f, 0.f, 0.f};
    if (!m.specular && pss_invertible(m) && pss_build(g, recs, m, sd, h)) {
        pe = guide_pdf_pss(&h, m, sd, F3{wi_eval[3 * i], wi_eval[3 * i + 1], wi_eval[3 * i + 2]});
        w = pss_sample_u(&h, m, sd, u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]);
        ps = guide_pdf_pss(&h, m, sd, w);
    }
    wi_out[3 * i] = w.x;
    wi_out[3 * i + 1] = w.y;
    wi_out[3 * i + 2] = w.z;
    pdf_sample[i] = ps;
    pdf_eval[i] = pe;
}

// RayTracer::computeDirect (Renderer.h:269-354) for a non-specular surface: picks a
// light, samples it and returns the unoccluded contribution and the shadow ray
// (false when the contribution is zero before visibility is known). In guided mode
// the BSDF density in the MIS weight is the one-sample mixture density.
// Guide density of MODE 1 (quadtree mixture of the nearest cells), MODE 2 (the
// learned vMF mixture of this queue slot) or MODE 3 (the PSS histogram).
struct PssRef {
    const PssHist* h;
    const MaterialRec* m;
    const SD* sd;
};
template <int MODE>
__device__ __forceinline__ float guide_pdf(const GuideView& g, const uint32_t* cells, const float4* mx, F3 wi,
                                           const PssRef& ps) {
    if (MODE == 1) return guide_pdf_qt(g, cells, wi);
    if (MODE == 3) return guide_pdf_pss(ps.h, *ps.m, *ps.sd, wi);
    return mix_pdf(mx, wi);
}
template <int MODE>
__device__ __forceinline__ F3 guide_sample(const GuideView& g, const uint32_t* cells, const float4* mx, uint32_t key,
                                           uint32_t& dim, const PssRef& ps) {
    if (MODE == 1) return guide_sample_qt(g, cells, key, dim);
    if (MODE == 3) return guide_sample_pss(ps.h, *ps.m, *ps.sd, key, dim);
    float u0 = keyed_rand(key, dim++);
    float u1 = keyed_rand(key, dim++);
    float u2 = keyed_rand(key, dim++);
    return mix_sample(mx, u0, u1, u2);
}

template <int MODE>
__device__ __forceinline__ bool next_event(const DevScene& s, const MaterialRec& m, const SD& sd, uint32_t key,
                                           uint32_t& dim, const GuideView& g, const uint32_t* cells,
                                           const float4* mx, bool guided, float ag, uint32_t env_vis, C3& contrib,
                                           F3& sh_o, F3& sh_d, float& sh_t, const PssRef& pr) {
    contrib = c3(0.f, 0.f, 0.f);
    float pmf = 1.f / (float)s.n_lights;
    uint32_t nl = (uint32_t)s.n_lights;
    float fl = floorf(keyed_rand(key, dim++) * (float)nl);
    uint32_t li = min(max(0u, (uint32_t)fl), nl - 1u);
    const LightRec& lr = s.lights[li];
    if (lr.kind == 0) {
        float r1 = keyed_rand(key, dim++), r2 = keyed_rand(key, dim++);
        float sr = sqrtf(r1);
        float ta = 1.f - sr, tb = r2 * sr, tg = (1.f - ta) - tb;
        float pdf_l = 1.f / lr.area;
        F3 pl = (ld3(lr.p0) * ta + ld3(lr.p1) * tb) + ld3(lr.p2) * tg;
        if (pdf_l <= 0.f) return false;
        F3 tol = pl - sd.x;
        F3 wi = normalize(tol);
        float sg = (dot(wi, sd.gn) >= 0.f) ? 1.f : -1.f;
        F3 off = sd.x + sd.gn * (kEps * sg);
        float den = length_sq(tol);
        if (den < kEps) return false;
        float ctl = smax(dot(wi, sd.sn), 0.f);
        float ctp = smax(dot(-wi, ld3(lr.gn)), 0.f);
        float G = (ctl * ctp) / den;
        float pdf_b = 0.f;
        C3 f = bsdf_eval(m, sd, wi, pdf_b);
        if (MODE != 0 && guided) pdf_b = (1.f - ag) * pdf_b + ag * guide_pdf<MODE>(g, cells, mx, wi, pr);
        float pal = pdf_l * pmf;
        float pab = (pdf_b * ctp) / den;
        float w = power_heuristic(pal, pab);
        contrib = (((c3(lr.emission[0], lr.emission[1], lr.emission[2]) * f) * G) * w) / pal;
        make_visible_ray(off, pl, sh_o, sh_d, sh_t);
        return true;
    }
    C3 le;
    float pdf_l = 0.f;
    F3 dl = env_sample(s, key, dim, le, pdf_l);
    if (pdf_l <= 0.f) return false;
    F3 wi = normalize(dl);
    float sg = (dot(wi, sd.gn) >= 0.f) ? 1.f : -1.f;
    F3 off = sd.x + sd.gn * (kEps * sg);
    // Renderer.h:331-332 aims the shadow ray at the scene-bound point in direction wi
    // from the centre, which is off the light direction for points away from the
    // centre; env_vis 1 traces along wi out of the bounding sphere instead.
    const F3 cen{s.centre[0], s.centre[1], s.centre[2]};
    F3 far = env_vis ? off + wi * (length(off - cen) + s.radius) : cen + wi * s.radius;
    float ctl = smax(dot(wi, sd.sn), 0.f);
    if (ctl <= 0.f) return false;
    float pdf_b = 0.f;
    C3 f = bsdf_eval(m, sd, wi, pdf_b);
    if (MODE != 0 && guided) pdf_b = (1.f - ag) * pdf_b + ag * guide_pdf<MODE>(g, cells, mx, wi, pr);
    float pal = pdf_l * pmf;
    float w = power_heuristic(pal, pdf_b);
    contrib = (((le * f) * ctl) * w) / pal;
    make_visible_ray(off, far, sh_o, sh_d, sh_t);
    return true;
}

// One bounce of RayTracer::pathTraceRecursive + computeDirect (Renderer.h:269-354, 531-613).
// MODE 0: BSDF sampling as the reference. MODE 1: one-sample MIS between the BSDF
// and the quadtree guide (probability guide_prob); MODE 2: between the BSDF and the
// learned vMF mixture (probability c, learned per point, at most guide_prob: the
// weight of a BSDF-sampled direction is then at most 1 / (1 - guide_prob) times the
// plain BSDF-sampling weight where the guide misses the radiance). The combined
// density is used in every MIS weight so that the estimator stays unbiased.
template <int MODE, bool MULTI>
__global__ __launch_bounds__(kShadeBlock, 3) void k_shade(DevScene s0, PathState ps, PassParams p, GuideCtx gc,
                                                       int parity, int depth) {
    __shared__ uint32_t app[2 * (kShadeBlock / 64) + 2];
    uint32_t n = ps.counters[parity];
    if (blockIdx.x == 0 && threadIdx.x == 0) ps.counters[4] = 0;  // shadow-ray claims
    const uint32_t* qin = ps.queue[parity];
    uint32_t* qout = ps.queue[parity ^ 1];
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        uint32_t i = base + threadIdx.x;
        bool active = i < n;
        bool want_shadow = false, want_ray = false;
        F3 sh_o{0.f, 0.f, 0.f}, sh_d{0.f, 0.f, 0.f}, nr_o{0.f, 0.f, 0.f}, nr_d{0.f, 0.f, 0.f};
        float sh_t = 0.f;
        C3 sh_c = c3(0.f, 0.f, 0.f);
        uint32_t path = 0;
        if (active) {
            path = qin[i];
            const DevScene& s = scene_of<MULTI>(s0, ps, path);
            float4 h = ps.hit[i];
            uint32_t hid = __float_as_uint(h.x);
            float ht = h.y, hu = h.z, hv = h.w;
            float4 ro = ps.ray_o[parity][i], rdv = ps.ray_d[parity][i];
            F3 rd{rdv.x, rdv.y, rdv.z};
            float4 b4 = ps.beta[path];
            C3 beta = c3(b4.x, b4.y, b4.z);
            float prev_pdf = b4.w;
            uint4 inf = ps.info[path];
            uint32_t key = inf.x, dim = inf.y;
            bool prev_spec = (inf.z & 1u) != 0u;
            float4 L4 = ps.radiance[path];
            C3 L = c3(L4.x, L4.y, L4.z);
            bool alive = false;
            if (hid == kMiss) {
                // background (Renderer.h:596-612)
                C3 bg = s.has_env ? env_eval(s, rd) : c3(0.f, 0.f, 0.f);
                if (depth == 0 || prev_spec) {
                    L = L + beta * bg;
                } else if (!(lum(bg) < 1e-8f)) {
                    float pmf = 1.f / (float)s.n_lights;
                    float pdf_l = s.has_env ? env_pdf(s, rd) : 0.f;
                    if (!(pmf <= 0.f || pdf_l <= 0.f)) {
                        float w = power_heuristic(prev_pdf, pmf * pdf_l);
                        L = L + (beta * bg) * w;
                    }
                }
            } else {
                // Scene::calculateShadingData (Scene.h:152-176)
                const TriShade& tsh = s.shade[hid];
                const MaterialRec& m = s.mats[tsh.mat];
                SD sd;
                F3 o{ro.x, ro.y, ro.z};
                sd.x = o + rd * ht;
                sd.gn = ld3(tsh.gn);
                float ha = 1.f - (hu + hv);
                sd.sn = normalize((ld3(tsh.n0) * ha + ld3(tsh.n1) * hu) + ld3(tsh.n2) * hv);
                sd.wo = -rd;
                if (m.two_sided) {
                    if (dot(sd.wo, sd.sn) < 0.f) sd.sn = -sd.sn;
                    if (dot(sd.wo, sd.gn) < 0.f) sd.gn = -sd.gn;
                }
                sd.fr = frame_from(sd.sn);
                sd.alb = material_albedo(s, m, hid, ha, hu, hv);
                if (m.is_light) {
                    C3 e = c3(m.emission[0], m.emission[1], m.emission[2]);
                    if (depth == 0 || prev_spec) {
                        L = L + beta * e;
                    } else {
                        float pmf = 1.f / (float)s.n_lights;
                        float pdf_l = 1.f / tsh.area;
                        float ctp = smax(dot(-rd, ld3(tsh.gn)), 0.f);
                        float d2 = ht * ht;
                        if (!(ctp <= 0.f) && !(d2 < kEps)) {
                            float pal = pdf_l * pmf;
                            float pab = (prev_pdf * ctp) / d2;
                            float w = power_heuristic(pab, pal);
                            L = L + (beta * e) * w;
                        }
                    }
                } else {
                    const uint32_t* cells = ((MODE == 1 || MODE == 3) && gc.cells) ? gc.cells + (size_t)i * kKnnMax
                                                                                    : nullptr;
                    const float4* mx = (MODE == 2 && gc.mix) ? gc.mix + (size_t)i * kMixF4 : nullptr;
                    float ag = gc.g.guide_prob;
                    bool guided = false;
                    if (MODE == 1) guided = !m.specular && cells != nullptr && cells[0] != 0xffffffffu;
                    PssHist ph;
                    const PssRef pr{&ph, &m, &sd};
                    if (MODE == 3)
                        guided = !m.specular && cells != nullptr && cells[0] != 0xffffffffu && pss_invertible(m) &&
                                 pss_build(gc.g, cells, m, sd, ph);
                    if (MODE == 2 && !m.specular && mx != nullptr) {
                        ag = fminf(mx[12].x, p.guide_cap);
                        guided = ag > 0.f;
                    }
                    // computeDirect (Renderer.h:269-354)
                    if (!m.specular && s.n_lights > 0) {
                        C3 contrib;
                        if (next_event<MODE>(s, m, sd, key, dim, gc.g, cells, mx, guided, ag, p.env_visibility, contrib,
                                             sh_o, sh_d, sh_t, pr)) {
                            // the shadow ray can only change the image when the contribution is nonzero
                            C3 direct = beta * contrib;
                            if (!is_zero(direct)) {
                                want_shadow = true;
                                sh_c = direct;
                            }
                        }
                    }
                    // Russian roulette (Renderer.h:561-565), depth cap (Renderer.h:568)
                    bool cont = true;
                    if (depth > (int)p.rr_depth) {
                        float rrp = smin(smax(kEps, lum(beta)), 1.f);
                        if (keyed_rand(key, dim++) < rrp) beta = beta / rrp;
                        else cont = false;
                    }
                    if (cont && depth == (int)p.max_depth) cont = false;
                    if (cont) {
                        float pdf = 0.f;
                        C3 f;
                        F3 wi;
                        if (guided) {
                            if (keyed_rand(key, dim++) < ag) {
                                wi = guide_sample<MODE>(gc.g, cells, mx, key, dim, pr);
                                f = bsdf_eval(m, sd, wi, pdf);
                                pdf = (1.f - ag) * pdf + ag * guide_pdf<MODE>(gc.g, cells, mx, wi, pr);
                            } else {
                                wi = bsdf_sample(m, sd, key, dim, f, pdf);
                                if (!(pdf <= 0.f))
                                    pdf = (1.f - ag) * pdf + ag * guide_pdf<MODE>(gc.g, cells, mx, wi, pr);
                            }
                        } else {
                            wi = bsdf_sample(m, sd, key, dim, f, pdf);
                        }
                        if (!(pdf <= 0.f)) {
                            float sg = (dot(wi, sd.gn) >= 0.f) ? 1.f : -1.f;
                            float ctw = fabsf(dot(wi, sd.sn));
                            if (!(ctw <= 0.f)) {
                                beta = ((beta * f) * ctw) / pdf;
                                bool ok = !(lum(beta) <= 0.f) && finite3(beta) &&
                                          !(beta.r < 0.f || beta.g < 0.f || beta.b < 0.f);
                                if (ok) {
                                    alive = true;
                                    nr_o = sd.x + sd.gn * (kEps * sg);
                                    nr_d = wi;
                                    prev_pdf = pdf;
                                    prev_spec = m.specular != 0;
                                }
                            }
                        }
                    }
                }
            }
            ps.radiance[path] = make_float4(L.r, L.g, L.b, 0.f);
            if (alive) {
                want_ray = true;
                ps.beta[path] = make_float4(beta.r, beta.g, beta.b, prev_pdf);
                ps.info[path] = make_uint4(key, dim, prev_spec ? 1u : 0u, inf.w);
            } else if (want_shadow) {
                ps.info[path] = make_uint4(key, dim, inf.z, inf.w);
            }
        }
        uint32_t sslot, rslot;
        block_append2<kShadeBlock>(ps.counters + 2, want_shadow, sslot, ps.counters + (parity ^ 1), want_ray, rslot, app);
        if (want_shadow) {
            ps.sh_o[sslot] = make_float4(sh_o.x, sh_o.y, sh_o.z, sh_t);
            ps.sh_d[sslot] = make_float4(sh_d.x, sh_d.y, sh_d.z, __uint_as_float(path));
            ps.sh_c[sslot] = make_float4(sh_c.r, sh_c.g, sh_c.b, 0.f);
        }
        if (want_ray) {
            qout[rslot] = path;
            ps.ray_o[parity ^ 1][rslot] = make_float4(nr_o.x, nr_o.y, nr_o.z, 0.f);
            ps.ray_d[parity ^ 1][rslot] = make_float4(nr_d.x, nr_d.y, nr_d.z, 0.f);
        }
    }
}

// Query order for the kNN of a bounce: a 24-bit Morton key of each shading point in
// the scene cube (slots past the queue, misses, lights and specular surfaces get the
// largest key), sorted so that the lanes of a wave search neighbouring points.
__device__ __forceinline__ uint32_t spread8(uint32_t x) {
    x &= 0xffu;
    x = (x | (x << 8)) & 0x0000F00Fu;
    x = (x | (x << 4)) & 0x000C30C3u;
    x = (x | (x << 2)) & 0x00249249u;
    return x;
}
__global__ void k_query_keys(DevScene s, PathState ps, float3 lo, float inv_ext, int parity, uint32_t capacity,
                             uint32_t* keys, uint32_t* vals, int shift) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= capacity) return;
    uint32_t key = 0xffffffu;
    if (i < ps.counters[parity]) {
        float4 h = ps.hit[i];
        uint32_t hid = __float_as_uint(h.x);
        if (hid != kMiss) {
            const MaterialRec& m = s.mats[s.shade[hid].mat];
            if (!m.specular && !m.is_light) {
                float4 ro = ps.ray_o[parity][i], rdv = ps.ray_d[parity][i];
                F3 x = F3{ro.x, ro.y, ro.z} + F3{rdv.x, rdv.y, rdv.z} * h.y;
                uint32_t qx = (uint32_t)min(max((int)((x.x - lo.x) * inv_ext * 256.f), 0), 255);
                uint32_t qy = (uint32_t)min(max((int)((x.y - lo.y) * inv_ext * 256.f), 0), 255);
                uint32_t qz = (uint32_t)min(max((int)((x.z - lo.z) * inv_ext * 256.f), 0), 255);
                key = (spread8(qx) << 2) | (spread8(qy) << 1) | spread8(qz);
            }
        }
    }
    keys[i] = key >> shift;
    vals[i] = i;
}

// Trace order of a ray queue: key = direction octant (3 bits) above a 20-bit Morton
// code of the origin in the scene cube (7/7/6 bits); slots past the queue get
// 2^24 - 1, which sorts them last, so the first `count` entries of the sorted
// permutation are exactly the queue. shadow = 1 keys the shadow queue.
__global__ void k_ray_keys(PathState ps, float3 lo, float inv_ext, int parity, int shadow, uint32_t capacity,
                           uint32_t* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= capacity) return;
    const uint32_t n = shadow ? ps.counters[2] : ps.counters[parity];
    uint32_t key = 0xffffffu;
    if (i < n) {
        float4 o = shadow ? ps.sh_o[i] : ps.ray_o[parity][i];
        float4 d = shadow ? ps.sh_d[i] : ps.ray_d[parity][i];
        uint32_t oct = (d.x < 0.f ? 4u : 0u) | (d.y < 0.f ? 2u : 0u) | (d.z < 0.f ? 1u : 0u);
        uint32_t qx = (uint32_t)min(max((int)((o.x - lo.x) * inv_ext * 128.f), 0), 127);
        uint32_t qy = (uint32_t)min(max((int)((o.y - lo.y) * inv_ext * 128.f), 0), 127);
        uint32_t qz = (uint32_t)min(max((int)((o.z - lo.z) * inv_ext * 64.f), 0), 63);
        uint32_t m = (spread8(qx) << 2) | (spread8(qy) << 1) | spread8(qz);
        key = (oct << 20) | (m & 0xfffffu);
    }
    keys[i] = key;
    vals[i] = i;
}

// Cells of the k nearest records for every shading point of this bounce, visited in
// `order` (a permutation of the queue capacity; slots past the queue are skipped).
constexpr int kGuideStack = 32;
template <int OUT, bool SEEDED, int STACK>
__global__ __launch_bounds__(kBlock) void k_guide_cells(DevScene s, PathState ps, GuideView g, uint32_t* cells,
                                                        const uint32_t* __restrict__ order, uint32_t capacity,
                                                        int parity) {
    constexpr int records_out = OUT;
    __shared__ int sn[(OUT == 2 ? 1 : STACK) * kBlock];
    __shared__ float sd_[(OUT == 2 ? 1 : STACK) * kBlock];
    uint32_t n = ps.counters[parity];
    const uint32_t m_end = order ? capacity : n;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < m_end; j += gridDim.x * kBlock) {
        const uint32_t i = order ? order[j] : j;
        if (i >= n) continue;
        float4 h = ps.hit[i];
        uint32_t hid = __float_as_uint(h.x);
        uint32_t* out = cells + (size_t)i * kKnnMax;
        bool ok = hid != kMiss && g.n_records > 0;
        if (ok) {
            const MaterialRec& m = s.mats[s.shade[hid].mat];
            ok = !m.specular && !m.is_light;
        }
        if (!ok) {
            out[0] = 0xffffffffu;
            continue;
        }
        float4 ro = ps.ray_o[parity][i], rdv = ps.ray_d[parity][i];
        F3 x = F3{ro.x, ro.y, ro.z} + F3{rdv.x, rdv.y, rdv.z} * h.y;
        if (records_out == 2) {  // approximate neighbours (the query's leaf list)
            nearby_records(g, x, out);
            continue;
        }
        float bd[kKnnMax];
        uint32_t bs[kKnnMax];
        // exact search; SEEDED prunes from the start with the bound of the query's leaf list
        knn_query<kKnnMax>(g, x, bd, bs, sn + threadIdx.x, sd_ + threadIdx.x, kBlock,
                           SEEDED ? knn_cap(g, x) : __int_as_float(0x7f800000));
        if (records_out) {  // the k nearest records themselves (sorted index), nearest first
            for (int j = 0; j < kKnnMax; ++j) out[j] = bs[j];
            continue;
        }
        // distinct cells with neighbour counts, nearest first
        int nu = 0;
        for (int j = 0; j < kKnnMax; ++j) {
            if (bs[j] == 0xffffffffu) break;
            uint32_t c = bs[j] / g.cell_size;
            int f = -1;
            for (int m = 0; m < nu; ++m)
                if ((out[m] & 0xffffffu) == c) { f = m; break; }
            if (f >= 0) out[f] += 1u << 24;
            else out[nu++] = c | (1u << 24);
        }
        if (nu < kKnnMax) out[nu] = 0xffffffffu;
    }
}

// Per point-BVH leaf, the kKnnMax records nearest the centroid of its records (exact
// knn_query), for the approximate neighbours of nearby_records.
__global__ __launch_bounds__(kBlock) void k_leaf_nbrs(GuideView g, uint32_t n_leaves, uint32_t* out) {
    __shared__ int sn[kGuideStack * kBlock];
    __shared__ float sd_[kGuideStack * kBlock];
    uint32_t l = blockIdx.x * kBlock + threadIdx.x;
    if (l >= n_leaves) return;
    uint32_t b = l * kLeafRecords, e = min(b + (uint32_t)kLeafRecords, g.n_records);
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (uint32_t s = b; s < e; ++s) {
        float4 p = g.rec_pos[s];
        cx += p.x; cy += p.y; cz += p.z;
    }
    float inv = 1.f / (float)(e - b);
    float bd[kKnnMax];
    uint32_t bs[kKnnMax];
    knn_query<kKnnMax>(g, F3{cx * inv, cy * inv, cz * inv}, bd, bs, sn + threadIdx.x, sd_ + threadIdx.x, kBlock);
    for (int j = 0; j < kKnnMax; ++j) out[(size_t)l * kKnnMax + j] = bs[j];
}

// One bounce of the record pass, RayTracer::guidedPathRecursive (Renderer.h:397-528):
// BSDF-sampled continuation with a fixed Russian-roulette probability 0.7 after
// depth rr_depth, one forward vertex {x, n, wi, f cos/(pdf rrp), emission,
// MIS-weighted emission, NEE, store} per bounce. With NEE == false the next-event
// estimate draws its random numbers as the reference does but is not traced (the
// reference's Li does not use it); with NEE == true a shadow ray is queued whose
// result lands in the vertex's NEE slot (the pixelColour recursion).
// MODE 2 samples the continuation by one-sample MIS between the BSDF and the learned
// mixture of the queue slot (probability c), with the combined density q in the
// vertex weight, the NEE weight and the next vertex's MIS weights; each vertex keeps
// q and the BSDF pdf of its direction as training data.
template <bool NEE, int MODE>
__global__ __launch_bounds__(kShadeBlock, 3) void k_shade_record(DevScene s, PathState ps, PassParams p, RecordState rs,
                                                              const float4* __restrict__ mix, int parity, int depth) {
    __shared__ uint32_t app[2 * (kShadeBlock / 64) + 2];
    uint32_t n = ps.counters[parity];
    const uint32_t* qin = ps.queue[parity];
    uint32_t* qout = ps.queue[parity ^ 1];
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        uint32_t i = base + threadIdx.x;
        bool active = i < n;
        bool want_ray = false, want_shadow = false;
        F3 nr_o{0.f, 0.f, 0.f}, nr_d{0.f, 0.f, 0.f}, sh_o{0.f, 0.f, 0.f}, sh_d{0.f, 0.f, 0.f};
        float sh_t = 0.f;
        C3 sh_c = c3(0.f, 0.f, 0.f);
        uint32_t path = 0;
        if (active) {
            path = qin[i];
            float4 h = ps.hit[i];
            uint32_t hid = __float_as_uint(h.x);
            float ht = h.y, hu = h.z, hv = h.w;
            float4 ro = ps.ray_o[parity][i], rdv = ps.ray_d[parity][i];
            F3 rd{rdv.x, rdv.y, rdv.z};
            uint4 inf = ps.info[path];
            uint32_t key = inf.x, dim = inf.y;
            float prev_pdf = ps.beta[path].w;
            bool prev_spec = (inf.z & 1u) != 0u;
            float4 vpos = make_float4(0.f, 0.f, 0.f, 0.f), vnrm = vpos, vwi = vpos, vwo = vpos, vw = vpos, ve = vpos,
                   vme = vpos;
            if (hid == kMiss) {
                // Renderer.h:501-527
                C3 bg = s.has_env ? env_eval(s, rd) : c3(0.f, 0.f, 0.f);
                ve = make_float4(bg.r, bg.g, bg.b, 0.f);
                if (depth == 0 || prev_spec) {
                    vme = ve;
                } else if (lum(bg) > 1e-8f) {
                    float pmf = 1.f / (float)s.n_lights;
                    float pdf_l = s.has_env ? env_pdf(s, rd) : 0.f;
                    if (pmf > 0.f && pdf_l > 0.f) {
                        float pal = pmf * pdf_l, pab = prev_pdf;
                        if (pal >= 0.f && pab >= 0.f) {
                            C3 me = bg * power_heuristic(pab, pal);
                            vme = make_float4(me.r, me.g, me.b, 0.f);
                        }
                    }
                }
            } else {
                const TriShade& tsh = s.shade[hid];
                const MaterialRec& m = s.mats[tsh.mat];
                SD sd;
                F3 o{ro.x, ro.y, ro.z};
                sd.x = o + rd * ht;
                sd.gn = ld3(tsh.gn);
                float ha = 1.f - (hu + hv);
                sd.sn = normalize((ld3(tsh.n0) * ha + ld3(tsh.n1) * hu) + ld3(tsh.n2) * hv);
                sd.wo = -rd;
                if (m.two_sided) {
                    if (dot(sd.wo, sd.sn) < 0.f) sd.sn = -sd.sn;
                    if (dot(sd.wo, sd.gn) < 0.f) sd.gn = -sd.gn;
                }
                sd.fr = frame_from(sd.sn);
                sd.alb = material_albedo(s, m, hid, ha, hu, hv);
                vpos = make_float4(sd.x.x, sd.x.y, sd.x.z, 0.f);
                vnrm = make_float4(sd.sn.x, sd.sn.y, sd.sn.z, 0.f);
                if (m.is_light) {
                    // Renderer.h:430-460
                    C3 e = c3(m.emission[0], m.emission[1], m.emission[2]);
                    ve = make_float4(e.r, e.g, e.b, 0.f);
                    if (depth == 0 || prev_spec) {
                        vme = ve;
                    } else {
                        float pmf = 1.f / (float)s.n_lights;
                        float pdf_l = 1.f / tsh.area;
                        if (pmf > 0.f && pdf_l > 0.f) {
                            float ctp = smax(dot(-rd, ld3(tsh.gn)), 0.f);
                            float d2 = ht * ht;
                            if (ctp > 0.f && d2 > kEps) {
                                float pal = pdf_l * pmf, pab = (prev_pdf * ctp) / d2;
                                if (pal >= 0.f && pab >= 0.f) {
                                    C3 me = e * power_heuristic(pab, pal);
                                    vme = make_float4(me.r, me.g, me.b, 0.f);
                                }
                            }
                        }
                    }
                } else {
                    const float4* mx = (MODE == 2 && mix) ? mix + (size_t)i * kMixF4 : nullptr;
                    const float ag = mx ? fminf(mx[12].x, p.guide_cap) : 0.f;
                    const bool guided = MODE == 2 && !m.specular && ag > 0.f;
                    if (!m.specular && s.n_lights > 0) {
                        if (NEE) {
                            C3 contrib;
                            GuideView none{};
                            if (next_event<MODE>(s, m, sd, key, dim, none, nullptr, mx, guided, ag, p.env_visibility,
                                                 contrib, sh_o, sh_d, sh_t, PssRef{nullptr, &m, &sd}) &&
                                !is_zero(contrib)) {
                                want_shadow = true;
                                sh_c = contrib;
                            }
                        } else {
                            uint32_t nl = (uint32_t)s.n_lights;
                            float fl = floorf(keyed_rand(key, dim++) * (float)nl);
                            uint32_t li = min(max(0u, (uint32_t)fl), nl - 1u);
                            dim += (s.lights[li].kind == 0) ? 2u : 4u;
                        }
                    }
                    bool cont = true;
                    float rr = 1.f;
                    if (depth > (int)p.rr_depth) {
                        if (keyed_rand(key, dim++) < 0.7f) rr = 0.7f;
                        else cont = false;
                    }
                    if (cont && depth == (int)p.max_depth) cont = false;
                    if (cont) {
                        float pdf = 0.f, pb = 0.f;
                        C3 f;
                        F3 wi;
                        if (guided) {
                            const bool from_guide = keyed_rand(key, dim++) < ag;
                            if (from_guide) {
                                float u0 = keyed_rand(key, dim++);
                                float u1 = keyed_rand(key, dim++);
                                float u2 = keyed_rand(key, dim++);
                                wi = mix_sample(mx, u0, u1, u2);
                                f = bsdf_eval(m, sd, wi, pb);
                            } else {
                                wi = bsdf_sample(m, sd, key, dim, f, pb);
                            }
                            // a direction the BSDF gives no value ends the path (zero contribution)
                            pdf = ((from_guide || pb > 0.f) && !is_zero(f)) ? (1.f - ag) * pb + ag * mix_pdf(mx, wi)
                                                                           : 0.f;
                        } else {
                            wi = bsdf_sample(m, sd, key, dim, f, pdf);
                            pb = pdf;
                        }
                        float ctw = fabsf(dot(wi, sd.sn));
                        if (!(pdf <= 0.f || ctw <= 0.f)) {
                            C3 w = (f * ctw) / (pdf * rr);
                            vwi = make_float4(wi.x, wi.y, wi.z, pdf);
                            vwo = make_float4(sd.wo.x, sd.wo.y, sd.wo.z, pb);
                            vw = make_float4(w.r, w.g, w.b, 0.f);
                            vpos.w = m.specular ? 0.f : 1.f;
                            float sg = (dot(wi, sd.gn) >= 0.f) ? 1.f : -1.f;
                            nr_o = sd.x + sd.gn * (kEps * sg);
                            nr_d = wi;
                            want_ray = true;
                            prev_pdf = pdf;
                            prev_spec = m.specular != 0;
                        }
                    }
                }
            }
            size_t o = (size_t)depth * rs.stride + path;
            rs.vpos[o] = vpos;
            rs.vnrm[o] = vnrm;
            rs.vwi[o] = vwi;
            rs.vwo[o] = vwo;
            rs.vw[o] = vw;
            rs.ve[o] = ve;
            if (NEE) {
                rs.vme[o] = vme;
                rs.vd[o] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            if (want_ray) {
                ps.info[path] = make_uint4(key, dim, prev_spec ? 1u : 0u, inf.w);
                ps.beta[path] = make_float4(1.f, 1.f, 1.f, prev_pdf);
            } else {
                rs.nv[path] = (uint32_t)depth + 1u;
            }
        }
        uint32_t sslot, rslot;
        block_append2<kShadeBlock>(ps.counters + 2, NEE && want_shadow, sslot, ps.counters + (parity ^ 1), want_ray,
                                   rslot, app);
        if (NEE && want_shadow) {
            ps.sh_o[sslot] = make_float4(sh_o.x, sh_o.y, sh_o.z, sh_t);
            ps.sh_d[sslot] = make_float4(sh_d.x, sh_d.y, sh_d.z, __uint_as_float(path));
            ps.sh_c[sslot] = make_float4(sh_c.r, sh_c.g, sh_c.b, __int_as_float(depth));
        }
        if (want_ray) {
            qout[rslot] = path;
            ps.ray_o[parity ^ 1][rslot] = make_float4(nr_o.x, nr_o.y, nr_o.z, 0.f);
            ps.ray_d[parity ^ 1][rslot] = make_float4(nr_d.x, nr_d.y, nr_d.z, 0.f);
        }
    }
}

// Shadow rays of the record pass: an unoccluded contribution goes to its vertex.
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_shadow_record(DevScene s, PathState ps, RecordState rs) {
    __shared__ int lds[STACK * kBlock];
    uint32_t n = ps.counters[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) ps.stats[1] += n;
    uint32_t cn = 0, ct = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float4 a = ps.sh_o[i], b = ps.sh_d[i];
        F3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
        if (traverse_visible<false>(s, o, d, a.w, lane_stack<STACK>(lds, ps), cn, ct)) {
            uint32_t path = __float_as_uint(b.w);
            float4 c = ps.sh_c[i];
            size_t off = (size_t)__float_as_int(c.w) * rs.stride + path;
            rs.vd[off] = make_float4(c.x, c.y, c.z, 0.f);
        }
    }
}

// Box-filter splat (Imaging.h:186-217; Renderer.h:697-703): a sample at (px, py)
// lands on pixel (floor(px+.5), floor(py+.5)) when it lies strictly within half a
// pixel; non-finite samples are dropped. Each film pixel gathers its (at most
// four) source paths in source-pixel order, so the sum is deterministic.
__global__ void k_resolve(PathState ps, PassParams p, float4* film) {
    const uint32_t n = p.width * p.height;
    const uint32_t ns = p.n_samples ? p.n_samples : 1u;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int X = (int)(i % p.width), Y = (int)(i / p.width);
    float4 acc = film[i];
    // samples in order, each as one pass of its own would add it
    for (uint32_t smp = 0; smp < ns; ++smp)
    for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
            int sx = X + dx, sy = Y + dy;
            if (sx < 0 || sy < 0) continue;
            uint32_t q = smp * n + (uint32_t)sy * p.width + (uint32_t)sx;
            float2 fp = ps.film_pos[q];
            int tx = (int)floorf(fp.x + 0.5f), ty = (int)floorf(fp.y + 0.5f);
            if (tx != X || ty != Y) continue;
            if (!(fabsf((float)tx - fp.x) < 0.5f && fabsf((float)ty - fp.y) < 0.5f)) continue;
            float4 L = ps.radiance[q];
            C3 c = c3(L.x, L.y, L.z);
            if (!finite3(c)) continue;
            C3 add = (c * 1.f) / 1.f;
            acc = make_float4(acc.x + add.r, acc.y + add.g, acc.z + add.b, acc.w);
        }
    film[i] = acc;
}

__global__ void k_clear_film(float4* film, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) film[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Plain streaming copy: the HBM reference point for achieved bandwidth and for
// scaling the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_copy(const float4* src, float4* dst, size_t n, int cu_count, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)cu_count * 8), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ launchers
// The LDS stack holds at most one entry per tree level, so a stack of the tree's
// maximum leaf depth never overflows; `stack` is that bound rounded up.
#define PG_STACK_DISPATCH(stack, CALL) \
    do {                                 \
        if ((stack) <= 16) { CALL(16); } \
        else if ((stack) <= 32) { CALL(32); } \
        else if ((stack) <= 48) { CALL(48); } \
        else { CALL(64); }               \
    } while (0)

hipError_t launch_trace_closest(const DevScene& s, const float* rays6, uint32_t n, uint32_t* id, float* t, float* a,
                                float* b, float* g, int stack, bool count, unsigned long long* stats,
                                hipStream_t stream) {
    if (n == 0) return hipSuccess;
    dim3 grid((n + kBlock - 1) / kBlock);
#define PG_CALL(S)                                                                                                  \
    if (count) hipLaunchKernelGGL((k_trace_closest<true, S>), grid, dim3(kBlock), 0, stream, s, rays6, n, id, t, a, b, g, stats); \
    else hipLaunchKernelGGL((k_trace_closest<false, S>), grid, dim3(kBlock), 0, stream, s, rays6, n, id, t, a, b, g, stats)
    PG_STACK_DISPATCH(stack, PG_CALL);
#undef PG_CALL
    return hipGetLastError();
}
hipError_t launch_trace_shadow(const DevScene& s, const float* rays6, const float* max_t, uint32_t n, uint8_t* vis,
                               int stack, hipStream_t stream) {
    if (n == 0) return hipSuccess;
#define PG_CALL(S) \
    hipLaunchKernelGGL((k_trace_shadow<S>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, s, rays6, max_t, n, vis)
    PG_STACK_DISPATCH(stack, PG_CALL);
#undef PG_CALL
    return hipGetLastError();
}
// One BSDF query per thread with given random numbers (parity aid): sample(), then
// evaluate() and PDF() at wi_eval, invert() of the sampled direction and the flags.
// sd: x, wo, sN, gN (12 floats; frame from sN); out: 15 floats per query.
__global__ void k_bsdf_query(MaterialRec m, const float* sd, const float* u, const float* we, uint32_t n, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = sd + 12 * (size_t)i;
    SD s;
    s.x = F3{q[0], q[1], q[2]};
    s.wo = F3{q[3], q[4], q[5]};
    s.sn = F3{q[6], q[7], q[8]};
    s.gn = F3{q[9], q[10], q[11]};
    s.fr = frame_from(s.sn);
    s.alb = c3(m.albedo[0], m.albedo[1], m.albedo[2]);
    ListRng r{{u[3 * i], u[3 * i + 1], u[3 * i + 2]}, 0};
    C3 f;
    float pdf = 0.f;
    F3 wi = bsdf_sample(m, s, r, f, pdf);
    float pe = 0.f;
    C3 fe = bsdf_eval(m, s, F3{we[3 * i], we[3 * i + 1], we[3 * i + 2]}, pe);
    float iu = 0.f, iv = 0.f;
    bool inv = bsdf_invert(m, s, wi, iu, iv);
    float* o = out + 15 * (size_t)i;
    o[0] = wi.x; o[1] = wi.y; o[2] = wi.z;
    o[3] = f.r; o[4] = f.g; o[5] = f.b; o[6] = pdf;
    o[7] = fe.r; o[8] = fe.g; o[9] = fe.b; o[10] = pe;
    o[11] = inv ? iu : 0.f; o[12] = inv ? iv : 0.f;
    o[13] = m.specular ? 1.f : 0.f;
    o[14] = m.two_sided ? 1.f : 0.f;
}
hipError_t launch_bsdf_query(const MaterialRec& m, const float* sd, const float* u, const float* we, uint32_t n,
                             float* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bsdf_query, dim3((n + 255) / 256), dim3(256), 0, stream, m, sd, u, we, n, out);
    return hipGetLastError();
}
// Closest-hit arrays (id, t, alpha, beta, gamma) to packed pg_hit records (5 words each).
__global__ void k_pack_hits(const uint32_t* id, const float* t, const float* a, const float* b, const float* g,
                            uint32_t n, uint32_t* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* o = out + 5 * (size_t)i;
    o[0] = id[i];
    o[1] = __float_as_uint(t[i]);
    o[2] = __float_as_uint(a[i]);
    o[3] = __float_as_uint(b[i]);
    o[4] = __float_as_uint(g[i]);
}
hipError_t launch_pack_hits(const uint32_t* id, const float* t, const float* a, const float* b, const float* g,
                            uint32_t n, uint32_t* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_hits, dim3((n + 255) / 256), dim3(256), 0, stream, id, t, a, b, g, n, out);
    return hipGetLastError();
}
hipError_t launch_generate_rays(const DevScene& s, const float* pxy, uint32_t n, float* rays6, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_generate_rays, dim3((n + 255) / 256), dim3(256), 0, stream, s, pxy, n, rays6);
    return hipGetLastError();
}
hipError_t launch_rng(uint32_t seed, uint32_t sample, const uint32_t* pixel, const uint32_t* dim, uint32_t n, float* out,
                      hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rng, dim3((n + 255) / 256), dim3(256), 0, stream, seed, sample, pixel, dim, n, out);
    return hipGetLastError();
}
hipError_t launch_raygen(const DevScene& s, const PathState& ps, const PassParams& p, hipStream_t stream) {
    uint32_t n = p.width * p.height * (p.n_samples ? p.n_samples : 1u);
    hipLaunchKernelGGL(k_raygen, dim3((n + 255) / 256), dim3(256), 0, stream, s, ps, p);
    return hipGetLastError();
}
hipError_t launch_extend(const DevScene& s, const PathState& ps, int parity, int lds_stack, bool count, uint32_t grid,
                         hipStream_t stream, bool multi, const uint32_t* order, bool dyn) {
    if (ps.spill && grid * (uint32_t)kBlock > ps.spill_lanes) grid = ps.spill_lanes / kBlock;
    if (dyn) {
#define PG_CALL(S)                                                                                                   \
    if (multi) hipLaunchKernelGGL((k_extend_dyn<false, S, true>), dim3(grid), dim3(kBlock), 0, stream, s, ps, parity); \
    else if (count) hipLaunchKernelGGL((k_extend_dyn<true, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps, parity); \
    else hipLaunchKernelGGL((k_extend_dyn<false, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps, parity)
        PG_STACK_DISPATCH(lds_stack, PG_CALL);
#undef PG_CALL
        return hipGetLastError();
    }
#define PG_CALL(S)                                                                                               \
    if (multi) hipLaunchKernelGGL((k_extend<false, S, true>), dim3(grid), dim3(kBlock), 0, stream, s, ps, parity, order); \
    else if (count) hipLaunchKernelGGL((k_extend<true, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps, parity, order); \
    else hipLaunchKernelGGL((k_extend<false, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps, parity, order)
    PG_STACK_DISPATCH(lds_stack, PG_CALL);
#undef PG_CALL
    return hipGetLastError();
}
hipError_t launch_shade(const DevScene& s, const PathState& ps, const PassParams& p, const GuideCtx& gc, int mode,
                        int parity, int depth, uint32_t grid, hipStream_t stream, bool multi) {
    if (multi) {
        if (mode == 1 || mode == 3) return hipErrorInvalidValue;  // batches: unguided or learned guide
        if (mode == 2)
            hipLaunchKernelGGL((k_shade<2, true>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, gc, parity, depth);
        else
            hipLaunchKernelGGL((k_shade<0, true>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, gc, parity, depth);
    } else if (mode == 1) {
        hipLaunchKernelGGL((k_shade<1, false>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, gc, parity, depth);
    } else if (mode == 3) {
        hipLaunchKernelGGL((k_shade<3, false>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, gc, parity, depth);
    } else if (mode == 2) {
        hipLaunchKernelGGL((k_shade<2, false>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, gc, parity, depth);
    } else {
        hipLaunchKernelGGL((k_shade<0, false>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, gc, parity, depth);
    }
    return hipGetLastError();
}
hipError_t launch_shadow(const DevScene& s, const PathState& ps, int lds_stack, bool count, uint32_t grid,
                         hipStream_t stream, bool multi, const uint32_t* order, bool dyn) {
    if (ps.spill && grid * (uint32_t)kBlock > ps.spill_lanes) grid = ps.spill_lanes / kBlock;
    if (dyn) {
#define PG_CALL(S)                                                                                          \
    if (multi) hipLaunchKernelGGL((k_shadow_dyn<false, S, true>), dim3(grid), dim3(kBlock), 0, stream, s, ps);  \
    else if (count) hipLaunchKernelGGL((k_shadow_dyn<true, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps); \
    else hipLaunchKernelGGL((k_shadow_dyn<false, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps)
        PG_STACK_DISPATCH(lds_stack, PG_CALL);
#undef PG_CALL
        return hipGetLastError();
    }
#define PG_CALL(S)                                                                                         \
    if (multi) hipLaunchKernelGGL((k_shadow<false, S, true>), dim3(grid), dim3(kBlock), 0, stream, s, ps, order);  \
    else if (count) hipLaunchKernelGGL((k_shadow<true, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps, order); \
    else hipLaunchKernelGGL((k_shadow<false, S, false>), dim3(grid), dim3(kBlock), 0, stream, s, ps, order)
    PG_STACK_DISPATCH(lds_stack, PG_CALL);
#undef PG_CALL
    return hipGetLastError();
}

// ------------------------------------------------------------------ batches of views
__device__ __forceinline__ uint32_t view_of(const BatchView* views, uint32_t n_views, uint32_t g) {
    uint32_t lo = 0, hi = n_views - 1;  // last view with pixel_offset <= g
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (views[mid].pixel_offset <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// k_raygen for the chunk [base, base + n) of the batch's global paths, ns samples per
// pixel: path i is sample (*sample + i / n) of chunk path i % n. The RNG key of a path
// is (seed, sample, pixel within its view), so each view's samples are the ones a
// single-view render of it draws.
__global__ void k_raygen_batch(PathState ps, uint32_t n_views, uint32_t base, uint32_t n, uint32_t ns,
                               uint32_t seed, const uint32_t* sample) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ps.counters[0] = n * ns;
        ps.counters[1] = 0;
        ps.counters[2] = 0;
        ps.counters[3] = 0;
    }
    if (i >= n * ns) return;
    const uint32_t g = base + i % n, smp = i / n;
    const uint32_t v = view_of(ps.bviews, n_views, g);
    const BatchView& bv = ps.bviews[v];
    const uint32_t pix = g - bv.pixel_offset;
    uint32_t x = pix % bv.width, y = pix / bv.width;
    uint32_t key = path_key(seed, *sample + smp, pix);
    float px = (float)x + keyed_rand(key, 0);
    float py = (float)y + keyed_rand(key, 1);
    F3 o, d;
    generate_ray(bv.cam, px, py, o, d);
    ps.ray_o[0][i] = make_float4(o.x, o.y, o.z, 0.f);
    ps.ray_d[0][i] = make_float4(d.x, d.y, d.z, 0.f);
    ps.queue[0][i] = i;
    ps.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
    ps.radiance[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    ps.film_pos[i] = make_float2(px, py);
    ps.info[i] = make_uint4(key, 2u, 0u, pix);
    ps.path_view[i] = v;
}

// k_resolve for the chunk's views (chunks hold whole views, so every source path of a
// pixel lies in the chunk); film is indexed by global pixel. The ns samples of a pass
// are added in sample order, as ns one-sample passes would add them.
__global__ void k_resolve_batch(PathState ps, uint32_t n_views, uint32_t base, uint32_t n, uint32_t ns,
                                float4* film) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = base + i;
    const BatchView& bv = ps.bviews[view_of(ps.bviews, n_views, g)];
    const uint32_t pix = g - bv.pixel_offset;
    const uint32_t first = bv.pixel_offset - base;  // chunk-local index of the view's pixel 0
    int X = (int)(pix % bv.width), Y = (int)(pix / bv.width);
    float4 acc = film[g];
    for (uint32_t smp = 0; smp < ns; ++smp)
    for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 0; ++dx) {
            int sx = X + dx, sy = Y + dy;
            if (sx < 0 || sy < 0) continue;
            uint32_t q = smp * n + first + (uint32_t)sy * bv.width + (uint32_t)sx;
            float2 fp = ps.film_pos[q];
            int tx = (int)floorf(fp.x + 0.5f), ty = (int)floorf(fp.y + 0.5f);
            if (tx != X || ty != Y) continue;
            if (!(fabsf((float)tx - fp.x) < 0.5f && fabsf((float)ty - fp.y) < 0.5f)) continue;
            float4 L = ps.radiance[q];
            C3 c = c3(L.x, L.y, L.z);
            if (!finite3(c)) continue;
            C3 add = (c * 1.f) / 1.f;
            acc = make_float4(acc.x + add.r, acc.y + add.g, acc.z + add.b, acc.w);
        }
    film[g] = acc;
}

__global__ void k_sample_advance(uint32_t* sample, uint32_t step) {
    if (threadIdx.x == 0) sample[0] += step;
}

hipError_t launch_raygen_batch(const PathState& ps, uint32_t n_views, uint32_t base, uint32_t n, uint32_t ns,
                               uint32_t seed, const uint32_t* sample, hipStream_t stream) {
    const uint32_t paths = n * ns;
    hipLaunchKernelGGL(k_raygen_batch, dim3((paths + 255) / 256), dim3(256), 0, stream, ps, n_views, base, n, ns,
                       seed, sample);
    return hipGetLastError();
}
hipError_t launch_resolve_batch(const PathState& ps, uint32_t n_views, uint32_t base, uint32_t n, uint32_t ns,
                                float4* film, hipStream_t stream) {
    hipLaunchKernelGGL(k_resolve_batch, dim3((n + 255) / 256), dim3(256), 0, stream, ps, n_views, base, n, ns, film);
    return hipGetLastError();
}
hipError_t launch_sample_advance(uint32_t* sample, uint32_t step, hipStream_t stream) {
    hipLaunchKernelGGL(k_sample_advance, dim3(1), dim3(64), 0, stream, sample, step);
    return hipGetLastError();
}

hipError_t launch_guide_cells(const DevScene& s, const PathState& ps, const GuideView& g, uint32_t* cells, int parity,
                              uint32_t grid, hipStream_t stream, const uint32_t* order, uint32_t capacity, int records_out) {
    // the search stack holds at most one entry per tree level: a 16-entry stack (half
    // the LDS) serves trees of up to 2^16 leaves
    const bool small = g.levels <= 16;
    const bool seeded = g.leaf_nbrs != nullptr;
#define PG_GC(OUT, SEED, ST) \
    hipLaunchKernelGGL((k_guide_cells<OUT, SEED, ST>), dim3(grid), dim3(kBlock), 0, stream, s, ps, g, cells, order, \
                       capacity, parity)
    if (records_out == 2) PG_GC(2, false, 16);
    else if (records_out == 1) {
        if (seeded) { if (small) PG_GC(1, true, 16); else PG_GC(1, true, 32); }
        else { if (small) PG_GC(1, false, 16); else PG_GC(1, false, 32); }
    } else {
        if (seeded) { if (small) PG_GC(0, true, 16); else PG_GC(0, true, 32); }
        else { if (small) PG_GC(0, false, 16); else PG_GC(0, false, 32); }
    }
#undef PG_GC
    return hipGetLastError();
}
hipError_t launch_leaf_nbrs(const GuideView& g, uint32_t n_leaves, uint32_t* out, hipStream_t stream) {
    if (n_leaves == 0) return hipSuccess;
    hipLaunchKernelGGL(k_leaf_nbrs, dim3((n_leaves + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, g, n_leaves, out);
    return hipGetLastError();
}
hipError_t launch_ray_keys(const PathState& ps, float3 lo, float inv_ext, int parity, int shadow, uint32_t capacity,
                           uint32_t* keys, uint32_t* vals, hipStream_t stream) {
    hipLaunchKernelGGL(k_ray_keys, dim3((capacity + 255) / 256), dim3(256), 0, stream, ps, lo, inv_ext, parity, shadow,
                       capacity, keys, vals);
    return hipGetLastError();
}
hipError_t launch_query_keys(const DevScene& s, const PathState& ps, float3 lo, float inv_ext, int parity,
                             uint32_t capacity, uint32_t* keys, uint32_t* vals, hipStream_t stream, int key_bits) {
    hipLaunchKernelGGL(k_query_keys, dim3((capacity + 255) / 256), dim3(256), 0, stream, s, ps, lo, inv_ext, parity,
                       capacity, keys, vals, 24 - key_bits);
    return hipGetLastError();
}
hipError_t launch_shade_record(const DevScene& s, const PathState& ps, const PassParams& p, const RecordState& rs,
                               const float4* mix, int parity, int depth, uint32_t grid, hipStream_t stream) {
    if (mix && rs.li_nee)
        hipLaunchKernelGGL((k_shade_record<true, 2>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, rs, mix, parity, depth);
    else if (mix)
        hipLaunchKernelGGL((k_shade_record<false, 2>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, rs, mix, parity, depth);
    else if (rs.li_nee)
        hipLaunchKernelGGL((k_shade_record<true, 0>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, rs, nullptr, parity, depth);
    else
        hipLaunchKernelGGL((k_shade_record<false, 0>), dim3(grid), dim3(kShadeBlock), 0, stream, s, ps, p, rs, nullptr, parity, depth);
    return hipGetLastError();
}
hipError_t launch_shadow_record(const DevScene& s, const PathState& ps, const RecordState& rs, int lds_stack, uint32_t grid,
                                hipStream_t stream) {
    if (ps.spill && grid * (uint32_t)kBlock > ps.spill_lanes) grid = ps.spill_lanes / kBlock;
#define PG_CALL(S) hipLaunchKernelGGL((k_shadow_record<S>), dim3(grid), dim3(kBlock), 0, stream, s, ps, rs)
    PG_STACK_DISPATCH(lds_stack, PG_CALL);
#undef PG_CALL
    return hipGetLastError();
}

int stack_entries(int max_depth) {
    if (max_depth <= 16) return 16;
    if (max_depth <= 32) return 32;
    if (max_depth <= 48) return 48;
    return 64;
}

// LDS entries of a render kernel's traversal stack: the whole depth without a spill
// buffer, else min(depth, lds_cap rounded up to 16) with the rest in global memory
// (lds_cap 0: kLdsStackEntries).
int lds_entries(int max_depth, uint32_t lds_cap, const PathState& ps) {
    const int whole = stack_entries(max_depth);
    if (!ps.spill) return whole;
    return std::min(whole, stack_entries(lds_cap ? (int)lds_cap : kLdsStackEntries));
}

// Persistent grid for the traversal kernels: as many 128-lane blocks per CU as
// the LDS stacks (160 KB per CU) allow, at most 16 (12 for the persistent-lane
// kernels, which keep their lanes busy with fewer waves and thrash L1 less).
uint32_t trace_grid(int cu_count, int lds_stack, bool dyn) {
    int per_cu = (160 * 1024) / (stack_entries(lds_stack) * kBlock * 4);
    if (per_cu > (dyn ? 12 : 16)) per_cu = dyn ? 12 : 16;
    if (per_cu < 1) per_cu = 1;
    return (uint32_t)(cu_count * per_cu);
}
hipError_t launch_pss_eval(const DevScene& s, const GuideView& g, uint32_t mat, const float* pos, const float* nrm,
                           const float* wo, const float* u3, const float* wi_eval, uint32_t n, float* wi_out,
                           float* pdf_sample, float* pdf_eval, hipStream_t stream) {
    hipLaunchKernelGGL(k_pss_eval, dim3((n + 127) / 128), dim3(128), 0, stream, s, g, mat, pos, nrm, wo, u3, wi_eval, n,
                       wi_out, pdf_sample, pdf_eval);
    return hipGetLastError();
}
hipError_t launch_resolve(const PathState& ps, const PassParams& p, float4* film, hipStream_t stream) {
    uint32_t n = p.width * p.height;
    hipLaunchKernelGGL(k_resolve, dim3((n + 255) / 256), dim3(256), 0, stream, ps, p, film);
    return hipGetLastError();
}
hipError_t launch_clear_film(float4* film, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_film, dim3((n + 255) / 256), dim3(256), 0, stream, film, n);
    return hipGetLastError();
}

}  // namespace pg
