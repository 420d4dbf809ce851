// kernels.h — device scene view and kernel launch declarations shared by the
// kernel translation unit and the C-ABI glue.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pg_tracer.h"
#include "pg_common.h"
#include "guide.h"

namespace pg {

// Read-only device scene, passed by value to every kernel.
struct DevScene {
    const BvhNode* nodes;       // slot layout: node i at nodes[i + 1]
    const TriIsect* tris;       // leaf order
    const TriShade* shade;      // by triangle ID
    const MaterialRec* mats;
    const LightRec* lights;
    const float4* env_tex;      // RGBA32F texels
    const float* env_pdf;       // uv-space density per texel
    const float* env_prob;      // alias acceptance
    const uint32_t* env_alias;  // alias target
    const float4* tex_texels;   // albedo textures, RGBA32F, concatenated
    const int4* tex_info;       // per texture: first texel, width, height
    const float2* uv;           // per triangle ID: (u, v) of its three vertices
    int32_t n_lights;
    int32_t env_w, env_h, has_env;
    int32_t finite_bounds;      // 1: every BVH node bound is finite
    float centre[3];
    float radius;
    CameraRec cam;
};

// One view of a batch: camera, resolution, scene, and its first global path index
// (= first film pixel); views are ordered by pixel_offset.
struct BatchView {
    CameraRec cam;
    uint32_t scene, width, height, pixel_offset;
};

// Structure-of-arrays path state for one pass (one sample per pixel).
struct PathState {
    float4* beta;      // throughput rgb, previous BSDF pdf
    float4* radiance;  // accumulated L rgb
    float2* film_pos;  // film sample position (px, py)
    uint4* info;       // rng key, rng dim, previous-specular flag, pixel
    // ray queues (double-buffered), indexed by queue slot
    uint32_t* queue[2];
    float4* ray_o[2];   // origin.xyz, unused
    float4* ray_d[2];   // dir.xyz, unused
    float4* hit;        // id bits, t, beta(u), gamma(v) by queue slot
    // shadow queue
    float4* sh_o;       // origin.xyz, maxT
    float4* sh_d;       // dir.xyz, path id bits
    float4* sh_c;       // contribution rgb
    uint32_t* counters; // [0],[1] ray queue sizes, [2] shadow queue size
    unsigned long long* stats;  // [0] closest rays, [1] shadow rays, [2] nodes, [3] tris, [4] sh nodes, [5] sh tris
    uint32_t capacity;
    // traversal stack entries beyond the LDS share: (64 - 16) per lane of the traversal
    // grid, lane-interleaved (spill_lanes lanes); nullptr when every entry fits in LDS
    int* spill;
    uint32_t spill_lanes;
    // batches of views over several scenes: view of each path, the scene table and
    // the view table (nullptr for single-view rendering)
    uint32_t* path_view;
    const DevScene* bscenes;
    const BatchView* bviews;
};

// Guide inputs of the shade kernel: the cache and, per queue slot, the cells of the
// k nearest records (kKnnMax entries, 0xffffffff when unavailable).
struct GuideCtx {
    GuideView g;
    const uint32_t* cells;
    const float4* mix;  // learned guide: kMixF4 float4 per queue slot, c = 0 where unguided
};

struct PassParams {
    uint32_t seed, sample;
    uint32_t width, height;
    uint32_t max_depth, rr_depth;
    int32_t integrator;
    uint32_t env_visibility;  // 0: Renderer.h:331-332 shadow target; 1: along wi
    uint32_t n_samples;       // samples per pixel traced together in this pass (0 = 1):
                              // path i is sample (sample + i / (width*height)) of pixel
                              // i % (width*height)
    float guide_cap;          // learned guide: its selection probability is capped here
};

// Launch helpers (defined in kernels.hip); all enqueue on `stream`.
hipError_t launch_trace_closest(const DevScene& s, const float* rays6, uint32_t n, uint32_t* id, float* t,
                                float* a, float* b, float* g, int stack, bool count,
                                unsigned long long* stats, hipStream_t stream);
hipError_t launch_trace_shadow(const DevScene& s, const float* rays6, const float* max_t, uint32_t n,
                               uint8_t* vis, int stack, hipStream_t stream);
hipError_t launch_pack_hits(const uint32_t* id, const float* t, const float* a, const float* b, const float* g,
                            uint32_t n, uint32_t* out, hipStream_t stream);
hipError_t launch_generate_rays(const DevScene& s, const float* pxy, uint32_t n, float* rays6,
                                hipStream_t stream);
hipError_t launch_rng(uint32_t seed, uint32_t sample, const uint32_t* pixel, const uint32_t* dim, uint32_t n,
                      float* out, hipStream_t stream);
// BSDF sample / evaluate / PDF / invert with given random numbers (15 floats out per query).
hipError_t launch_bsdf_query(const MaterialRec& m, const float* sd, const float* u, const float* we, uint32_t n,
                             float* out, hipStream_t stream);

hipError_t launch_raygen(const DevScene& s, const PathState& ps, const PassParams& p, hipStream_t stream);
// order (nullable): trace order of the queue (a sorted permutation from launch_ray_keys)
hipError_t launch_extend(const DevScene& s, const PathState& ps, int parity, int lds_stack, bool count,
                         uint32_t grid, hipStream_t stream, bool multi = false, const uint32_t* order = nullptr,
                         bool dyn = false);
hipError_t launch_shade(const DevScene& s, const PathState& ps, const PassParams& p, const GuideCtx& gc, int mode,
                        int parity, int depth, uint32_t grid, hipStream_t stream, bool multi = false);
hipError_t launch_shadow(const DevScene& s, const PathState& ps, int lds_stack, bool count, uint32_t grid,
                         hipStream_t stream, bool multi = false, const uint32_t* order = nullptr, bool dyn = false);
hipError_t launch_ray_keys(const PathState& ps, float3 lo, float inv_ext, int parity, int shadow, uint32_t capacity,
                           uint32_t* keys, uint32_t* vals, hipStream_t stream);
// Batches: ray generation for the chunk of global paths [base, base + n), ns samples
// per pixel (first sample index read from *sample), film splat of the chunk's views
// into film[global pixel], and the per-pass sample advance by ns, all graph-capturable
// (no host values change between passes).
hipError_t launch_raygen_batch(const PathState& ps, uint32_t n_views, uint32_t base, uint32_t n, uint32_t ns,
                               uint32_t seed, const uint32_t* sample, hipStream_t stream);
hipError_t launch_resolve_batch(const PathState& ps, uint32_t n_views, uint32_t base, uint32_t n, uint32_t ns,
                                float4* film, hipStream_t stream);
hipError_t launch_sample_advance(uint32_t* sample, uint32_t step, hipStream_t stream);
hipError_t launch_resolve(const PathState& ps, const PassParams& p, float4* film, hipStream_t stream);
// The primary-sample-space guide at n explicit points of material `mat` (3 floats per
// point for pos, nrm, wo, u3, wi_eval and wi_out).
hipError_t launch_pss_eval(const DevScene& s, const GuideView& g, uint32_t mat, const float* pos, const float* nrm,
                           const float* wo, const float* u3, const float* wi_eval, uint32_t n, float* wi_out,
                           float* pdf_sample, float* pdf_eval, hipStream_t stream);
hipError_t launch_clear_film(float4* film, uint32_t n, hipStream_t stream);
hipError_t launch_copy(const float4* src, float4* dst, size_t n, int cu_count, hipStream_t stream);
struct RecordState;
// kNN cells per shading point; with `order` (a sorted permutation of the queue
// capacity from launch_query_keys + radix sort) waves visit neighbouring points.
hipError_t launch_leaf_nbrs(const GuideView& g, uint32_t n_leaves, uint32_t* out, hipStream_t stream);
// records_out: 0 distinct cells of the k nearest records with counts (quadtree guide),
// 1 the k nearest records, 2 the precomputed neighbours of the query's Morton leaf (PSS guide).
hipError_t launch_guide_cells(const DevScene& s, const PathState& ps, const GuideView& g, uint32_t* cells, int parity,
                              uint32_t grid, hipStream_t stream, const uint32_t* order = nullptr,
                              uint32_t capacity = 0, int records_out = 0);
// key_bits < 24 keeps the top key_bits of the 24-bit Morton code (a coarser order that
// sorts in fewer radix passes)
hipError_t launch_query_keys(const DevScene& s, const PathState& ps, float3 lo, float inv_ext, int parity,
                             uint32_t capacity, uint32_t* keys, uint32_t* vals, hipStream_t stream, int key_bits = 24);
// mix != nullptr: learned-guide record pass (one-sample MIS with the mixtures of k_mlp_infer)
hipError_t launch_shade_record(const DevScene& s, const PathState& ps, const PassParams& p, const RecordState& rs,
                               const float4* mix, int parity, int depth, uint32_t grid, hipStream_t stream);
hipError_t launch_shadow_record(const DevScene& s, const PathState& ps, const RecordState& rs, int lds_stack, uint32_t grid,
                                hipStream_t stream);
struct MlpView;
// order (nullable): the queue slots in Morton order of their shading points (a sorted
// permutation whose first counters[parity] entries are the queue); the eighth of them
// that XCD x evaluates stays in that XCD's L2 (see k_mlp_infer)
hipError_t launch_mlp_infer(const DevScene& s, const PathState& ps, const MlpView& v, int parity, float4* mix,
                            uint32_t grid, hipStream_t stream, bool multi = false, uint32_t scene = 0,
                            const uint32_t* order = nullptr);
// Batches: the queue slots grouped by scene (cnt, off, cur: n_scenes + 1 entries each;
// perm: queue capacity + 128 * n_scenes; n_tiles: 1) so that one launch runs every
// scene's network (guides[s].wmirror == nullptr: scene s is unguided).
struct GroupBuffers {
    uint32_t *cnt, *off, *cur, *perm, *n_tiles;
};
hipError_t launch_mlp_infer_grouped(const PathState& ps, int parity, const MlpView* guides, uint32_t n_scenes,
                                    const GroupBuffers& gb, float4* mix, uint32_t grid, hipStream_t stream);
int stack_entries(int max_depth);
// Shading kernels (k_shade, k_shade_record): 256-lane blocks, whose queue appends are
// aggregated over the block; launched at 3 blocks per CU unless set.
constexpr int kShadeBlock = 256;
constexpr uint32_t kShadeBlocksPerCu = 3;
// Render kernels keep the first lds_entries() traversal stack entries of a lane in LDS
// (by default kLdsStackEntries) and the rest in PathState::spill (at most kSpillEntries
// per lane); their launchers take that LDS count as `lds_stack`.
constexpr int kLdsStackEntries = 16;
constexpr int kSpillEntries = 64 - kLdsStackEntries;
int lds_entries(int max_depth, uint32_t lds_cap, const PathState& ps);
uint32_t trace_grid(int cu_count, int lds_stack, bool dyn);

}  // namespace pg
