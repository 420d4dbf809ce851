// batch.hip — batches of views over several scenes, replayed as a hipGraph per SPP,
// with an autotuner for queue size, launch grids and traversal stack depth.
//
// The reference renders one camera per RayTracer::render (Renderer.h:648-747); many
// small views rendered one after another are launch-bound on a GPU (a 128x128 view
// is 16K paths: a few microseconds of work per kernel). A batch puts all views' paths
// into one global index space (view-major), runs them through the same persistent
// grid-stride kernels as the single-view path, with the scene chosen per path, in
// view-aligned chunks, and captures the whole pass (every chunk, every bounce) once
// as a hipGraph: one graph launch per SPP.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pg_tracer.h"
#include "ctx_internal.h"
#include "host_prep.hpp"
#include "kernels.h"
#include "mlp.h"

using namespace pg;

namespace {
template <class T>
hipError_t balloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}
template <class T>
void bfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// A scene's learned guide in a batch: FP32 master parameters and hash grid as the
// caller trained them (pg_mlp_get_params), their BF16 mirrors, and the view.
struct SceneGuide {
    float* params = nullptr;
    uint16_t* mirror = nullptr;
    float* grid = nullptr;
    uint32_t* grid_bf = nullptr;
    MlpView view{};
    ~SceneGuide() { bfree(params); bfree(mirror); bfree(grid); bfree(grid_bf); }
};

struct SceneBuf {
    PreparedScene host;
    std::unique_ptr<SceneGuide> guide;
    BvhNode* nodes = nullptr;
    TriIsect* tris = nullptr;
    TriShade* shade = nullptr;
    MaterialRec* mats = nullptr;
    LightRec* lights = nullptr;
    float4* env_tex = nullptr;
    float* env_pdf = nullptr;
    float* env_prob = nullptr;
    uint32_t* env_alias = nullptr;
    pg_internal::TextureSet tex;
    DevScene dev{};
    ~SceneBuf() {
        bfree(nodes); bfree(tris); bfree(shade); bfree(mats); bfree(lights);
        bfree(env_tex); bfree(env_pdf); bfree(env_prob); bfree(env_alias);
        tex.release();
    }
    // Same device layout as pg_scene_upload.
    hipError_t upload() {
        const PreparedScene& h = host;
        hipError_t e;
#define PG_B(x) if ((e = (x)) != hipSuccess) return e
        PG_B(balloc(&nodes, h.node_slots.size()));
        // one zeroed record past the end: traversal reads 64 bytes at every triangle
        PG_B(balloc(&tris, h.tris.size() + 1));
        PG_B(hipMemset(tris + h.tris.size(), 0, sizeof(TriIsect)));
        PG_B(balloc(&shade, h.shade.size()));
        PG_B(balloc(&mats, h.mats.size()));
        PG_B(balloc(&lights, h.lights.size() ? h.lights.size() : 1));
        PG_B(hipMemcpy(nodes, h.node_slots.data(), h.node_slots.size() * sizeof(BvhNode), hipMemcpyHostToDevice));
        PG_B(hipMemcpy(tris, h.tris.data(), h.tris.size() * sizeof(TriIsect), hipMemcpyHostToDevice));
        PG_B(hipMemcpy(shade, h.shade.data(), h.shade.size() * sizeof(TriShade), hipMemcpyHostToDevice));
        PG_B(hipMemcpy(mats, h.mats.data(), h.mats.size() * sizeof(MaterialRec), hipMemcpyHostToDevice));
        if (!h.lights.empty())
            PG_B(hipMemcpy(lights, h.lights.data(), h.lights.size() * sizeof(LightRec), hipMemcpyHostToDevice));
        if (h.has_env) {
            size_t n = (size_t)h.env.w * h.env.h;
            PG_B(balloc(&env_tex, n));
            PG_B(balloc(&env_pdf, n));
            PG_B(balloc(&env_prob, n));
            PG_B(balloc(&env_alias, n));
            PG_B(hipMemcpy(env_tex, h.env.texels.data(), n * 16, hipMemcpyHostToDevice));
            PG_B(hipMemcpy(env_pdf, h.env.pdf.data(), n * 4, hipMemcpyHostToDevice));
            PG_B(hipMemcpy(env_prob, h.env.prob.data(), n * 4, hipMemcpyHostToDevice));
            PG_B(hipMemcpy(env_alias, h.env.alias.data(), n * 4, hipMemcpyHostToDevice));
        }
#undef PG_B
        DevScene& d = dev;
        d.nodes = nodes; d.tris = tris; d.shade = shade; d.mats = mats; d.lights = lights;
        d.env_tex = env_tex; d.env_pdf = env_pdf; d.env_prob = env_prob; d.env_alias = env_alias;
        d.n_lights = (int32_t)h.lights.size();
        d.env_w = h.has_env ? h.env.w : 0;
        d.env_h = h.has_env ? h.env.h : 0;
        d.has_env = h.has_env ? 1 : 0;
        d.finite_bounds = h.finite_bounds ? 1 : 0;
        d.centre[0] = h.centre[0]; d.centre[1] = h.centre[1]; d.centre[2] = h.centre[2];
        d.radius = h.radius;
        d.cam = h.camera;
        return hipSuccess;
    }
};
}  // namespace

struct pg_batch {
    pg_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<SceneBuf>> scenes;
    std::vector<BatchView> views;
    uint32_t total = 0;
    int max_depth = 0;
    DevScene* d_scenes = nullptr;
    BatchView* d_views = nullptr;
    bool dirty = true;  // tables, film, chunks and path state need rebuilding
    PathState ps{};
    uint32_t ps_cap = 0;
    uint32_t* d_counters = nullptr;
    unsigned long long* d_stats = nullptr;
    uint32_t* d_sample = nullptr;
    float4* film = nullptr;
    uint32_t film_cap = 0;
    uint32_t spp = 0;
    pg_batch_config cfg{0, 0, 0, 0, 1, 0, 0, 0};
    std::vector<std::pair<uint32_t, uint32_t>> chunks;  // (first global path, paths)
    std::vector<std::vector<uint32_t>> chunk_scenes;    // scenes with views in each chunk
    float4* mix = nullptr;                              // learned-guide mixtures by queue slot
    uint32_t mix_cap = 0;
    // learned guides of the scenes for the grouped inference (wmirror null: no guide)
    MlpView* d_guides = nullptr;
    uint32_t guides_cap = 0;
    GroupBuffers grp{};
    uint32_t grp_cap = 0, grp_scenes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_ok = false;
    pg_render_params gparams{};
    uint32_t graph_ns = 0;  // samples per pass of the captured graph
    size_t graph_nodes = 0;
    uint64_t paths = 0;
    double gpu_ms = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    pg_status fail(pg_status code, const std::string& m) { return pg_internal::fail(ctx, code, m); }
    pg_status hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return PG_OK;
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
    hipStream_t stream() const { return pg_internal::stream(ctx); }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_ok = false;
        graph_nodes = 0;
    }
    void free_paths() {
        bfree(ps.beta); bfree(ps.radiance); bfree(ps.film_pos); bfree(ps.info);
        for (int k = 0; k < 2; ++k) { bfree(ps.queue[k]); bfree(ps.ray_o[k]); bfree(ps.ray_d[k]); }
        bfree(ps.hit); bfree(ps.sh_o); bfree(ps.sh_d); bfree(ps.sh_c); bfree(ps.path_view); bfree(ps.spill);
        ps.spill_lanes = 0;
        ps_cap = 0;
    }
    int stack() const {
        int s = cfg.stack ? (int)cfg.stack : max_depth;
        return stack_entries(std::max(s, 1));
    }
    uint32_t samples_per_pass() const { return cfg.samples_per_pass ? cfg.samples_per_pass : 1u; }
};

#define PG_BTRY(b, expr, what)                        \
    do {                                              \
        pg_status _s = (b)->hip((expr), (what));      \
        if (_s != PG_OK) return _s;                   \
    } while (0)

namespace {

// Tables, film, view-aligned chunks and the path state for the current configuration.
pg_status prepare(pg_batch* b) {
    if (!b->dirty) return PG_OK;
    if (b->views.empty()) return b->fail(PG_ERR_STATE, "batch has no views");
    b->drop_graph();
    PG_BTRY(b, hipStreamSynchronize(b->stream()), "sync");
    // tables
    bfree(b->d_scenes);
    bfree(b->d_views);
    std::vector<DevScene> ds;
    for (auto& s : b->scenes) ds.push_back(s->dev);
    PG_BTRY(b, balloc(&b->d_scenes, ds.size()), "alloc scene table");
    PG_BTRY(b, balloc(&b->d_views, b->views.size()), "alloc view table");
    PG_BTRY(b, hipMemcpy(b->d_scenes, ds.data(), ds.size() * sizeof(DevScene), hipMemcpyHostToDevice), "scene table");
    PG_BTRY(b, hipMemcpy(b->d_views, b->views.data(), b->views.size() * sizeof(BatchView), hipMemcpyHostToDevice),
            "view table");
    // film (global pixel index), kept across reconfigurations of the same views
    if (b->film_cap < b->total) {
        b->film_cap = 0;
        bfree(b->film);
        PG_BTRY(b, balloc(&b->film, b->total), "alloc batch film");
        b->film_cap = b->total;
        PG_BTRY(b, launch_clear_film(b->film, b->total, b->stream()), "clear film");
        b->spp = 0;
        PG_BTRY(b, hipMemset(b->d_sample, 0, 4), "reset sample");
    }
    // chunks: whole views, greedily packed up to queue_paths
    b->chunks.clear();
    b->chunk_scenes.clear();
    const uint64_t q = b->cfg.queue_paths ? b->cfg.queue_paths : ~0ull;
    uint32_t base = 0, n = 0, largest = 0;
    std::vector<uint32_t> cs;
    for (const BatchView& v : b->views) {
        uint32_t px = v.width * v.height;
        if (n > 0 && (uint64_t)n + px > q) {
            b->chunks.push_back({base, n});
            b->chunk_scenes.push_back(cs);
            cs.clear();
            base += n;
            n = 0;
        }
        n += px;
        if (std::find(cs.begin(), cs.end(), v.scene) == cs.end()) cs.push_back(v.scene);
    }
    if (n) {
        b->chunks.push_back({base, n});
        b->chunk_scenes.push_back(cs);
    }
    for (auto& c : b->chunks) largest = std::max(largest, c.second);
    // path state for the largest chunk times the samples traced together per pass
    const uint64_t need = (uint64_t)largest * b->samples_per_pass();
    if (b->samples_per_pass() > 1 && need > (1ull << 26))
        return b->fail(PG_ERR_INVALID, "samples_per_pass x chunk paths above 2^26 (lower one of them)");
    if (b->ps_cap < need) {
        b->ps_cap = 0;  // stays 0 until every buffer below is allocated
        b->free_paths();
        PathState& p = b->ps;
        const uint32_t m = (uint32_t)need;
        PG_BTRY(b, balloc(&p.beta, m), "alloc batch paths");
        PG_BTRY(b, balloc(&p.radiance, m), "alloc batch paths");
        PG_BTRY(b, balloc(&p.film_pos, m), "alloc batch paths");
        PG_BTRY(b, balloc(&p.info, m), "alloc batch paths");
        for (int k = 0; k < 2; ++k) {
            PG_BTRY(b, balloc(&p.queue[k], m), "alloc batch queues");
            PG_BTRY(b, balloc(&p.ray_o[k], m), "alloc batch queues");
            PG_BTRY(b, balloc(&p.ray_d[k], m), "alloc batch queues");
        }
        PG_BTRY(b, balloc(&p.hit, m), "alloc batch hits");
        PG_BTRY(b, balloc(&p.sh_o, m), "alloc batch shadow");
        PG_BTRY(b, balloc(&p.sh_d, m), "alloc batch shadow");
        PG_BTRY(b, balloc(&p.sh_c, m), "alloc batch shadow");
        PG_BTRY(b, balloc(&p.path_view, m), "alloc batch views");
        p.spill_lanes = (uint32_t)pg_internal::cu_count(b->ctx) * 32u * 128u;
        PG_BTRY(b, balloc(&p.spill, (size_t)p.spill_lanes * kSpillEntries), "alloc traversal stacks");
        p.capacity = m;
        b->ps_cap = m;
    }
    b->ps.counters = b->d_counters;
    b->ps.stats = b->d_stats;
    b->ps.bscenes = b->d_scenes;
    b->ps.bviews = b->d_views;
    b->dirty = false;
    return PG_OK;
}

// One pass (ns samples per pixel of every view) on the batch stream.
hipError_t enqueue_pass(pg_batch* b, const pg_render_params& prm, uint32_t ns) {
    const int cu = pg_internal::cu_count(b->ctx);
    const int stack = lds_entries(b->stack(), b->cfg.stack_lds, b->ps);
    const uint32_t tgrid =
        b->cfg.trace_blocks_per_cu ? (uint32_t)cu * b->cfg.trace_blocks_per_cu : trace_grid(cu, stack, b->cfg.dyn_fetch != 0);
    const uint32_t sgrid = (uint32_t)cu * (b->cfg.shade_blocks_per_cu ? b->cfg.shade_blocks_per_cu : kShadeBlocksPerCu);
    const DevScene& s0 = b->scenes[0]->dev;  // unused by the batch kernels' per-path scenes
    const uint32_t nv = (uint32_t)b->views.size();
    const bool learned = prm.integrator == PG_INTEGRATOR_GUIDED_MLP;
    GuideCtx gc{};
    if (learned) gc.mix = b->mix;
    hipStream_t st = b->stream();
    hipError_t e;
    for (size_t ci = 0; ci < b->chunks.size(); ++ci) {
        const auto& ch = b->chunks[ci];
        PassParams pp{prm.seed,         0u, 0u, 0u, prm.max_depth, prm.rr_depth, prm.integrator, prm.env_visibility,
                      1u,               prm.guide_prob};
        if ((e = launch_raygen_batch(b->ps, nv, ch.first, ch.second, ns, prm.seed, b->d_sample, st)) != hipSuccess)
            return e;
        for (uint32_t depth = 0; depth <= prm.max_depth; ++depth) {
            int parity = (int)(depth & 1u);
            if ((e = launch_extend(s0, b->ps, parity, stack, false, tgrid, st, true, nullptr, b->cfg.dyn_fetch != 0)) !=
                hipSuccess)
                return e;
            // guide_depth d: the network guides the bounces below depth d only
            const bool guide_here = learned && (prm.guide_depth == 0 || depth < prm.guide_depth);
            if (guide_here) {
                // one launch runs each scene's network over that scene's slots (grouped
                // by scene); the slots of scenes without a network become unguided
                if ((e = launch_mlp_infer_grouped(b->ps, parity, b->d_guides, (uint32_t)b->scenes.size(), b->grp,
                                                  b->mix, (uint32_t)cu * 3, st)) != hipSuccess)
                    return e;
            }
            if ((e = launch_shade(s0, b->ps, pp, gc, guide_here ? 2 : 0, parity, (int)depth, sgrid, st, true)) !=
                hipSuccess)
                return e;
            if ((e = launch_shadow(s0, b->ps, stack, false, tgrid, st, true, nullptr, b->cfg.dyn_fetch != 0)) !=
                hipSuccess)
                return e;
        }
        if ((e = launch_resolve_batch(b->ps, nv, ch.first, ch.second, ns, b->film, st)) != hipSuccess) return e;
    }
    return launch_sample_advance(b->d_sample, ns, st);
}

bool same_params(const pg_render_params& a, const pg_render_params& b) {
    // every field enqueue_pass bakes into the captured launches
    return a.seed == b.seed && a.max_depth == b.max_depth && a.rr_depth == b.rr_depth &&
           a.integrator == b.integrator && a.env_visibility == b.env_visibility && a.guide_prob == b.guide_prob &&
           a.guide_depth == b.guide_depth;
}

// Parameters a batch renders with: unguided, or the learned guides of its scenes.
pg_status check_params(pg_batch* b, const pg_render_params& prm) {
    if (prm.integrator != PG_INTEGRATOR_PATH && prm.integrator != PG_INTEGRATOR_GUIDED_MLP)
        return b->fail(PG_ERR_INVALID, "batches render unguided or with the scenes' learned guides");
    if (prm.integrator == PG_INTEGRATOR_GUIDED_MLP) {
        bool any = false;
        for (auto& s : b->scenes) any = any || s->guide != nullptr;
        if (!any) return b->fail(PG_ERR_STATE, "no scene of the batch has a learned guide (pg_batch_set_guide)");
    }
    if (prm.max_depth > 64) return b->fail(PG_ERR_INVALID, "max_depth > 64");
    if (prm.env_visibility > 1u) return b->fail(PG_ERR_INVALID, "env_visibility must be 0 or 1");
    if (!(prm.guide_prob >= 0.f && prm.guide_prob <= 1.f)) return b->fail(PG_ERR_INVALID, "guide_prob outside [0,1]");
    return PG_OK;
}

pg_status render(pg_batch* b, const pg_render_params& prm, uint32_t n_spp) {
    pg_status st = prepare(b);
    if (st != PG_OK) return st;
    hipStream_t s = b->stream();
    if (prm.integrator == PG_INTEGRATOR_GUIDED_MLP && b->mix_cap < b->ps_cap) {
        b->drop_graph();
        PG_BTRY(b, hipStreamSynchronize(s), "sync");
        b->mix_cap = 0;
        bfree(b->mix);
        PG_BTRY(b, balloc(&b->mix, (size_t)b->ps_cap * kMixF4), "alloc guide mixtures");
        b->mix_cap = b->ps_cap;
    }
    if (prm.integrator == PG_INTEGRATOR_GUIDED_MLP) {
        // the scenes' guide table and the grouping buffers of the one-launch inference
        const uint32_t ns_ = (uint32_t)b->scenes.size();
        const uint32_t cap = b->ps_cap + 128u * ns_;
        if (b->grp_cap < cap || b->grp_scenes < ns_ || b->guides_cap < ns_) {
            b->drop_graph();
            PG_BTRY(b, hipStreamSynchronize(s), "sync");
            b->grp_cap = b->grp_scenes = b->guides_cap = 0;
            bfree(b->grp.cnt); bfree(b->grp.off); bfree(b->grp.cur); bfree(b->grp.perm); bfree(b->grp.n_tiles);
            bfree(b->d_guides);
            PG_BTRY(b, balloc(&b->grp.cnt, ns_ + 1), "alloc grouping");
            PG_BTRY(b, balloc(&b->grp.off, ns_ + 1), "alloc grouping");
            PG_BTRY(b, balloc(&b->grp.cur, ns_ + 1), "alloc grouping");
            PG_BTRY(b, balloc(&b->grp.perm, cap), "alloc grouping");
            PG_BTRY(b, balloc(&b->grp.n_tiles, 1), "alloc grouping");
            PG_BTRY(b, balloc(&b->d_guides, ns_), "alloc guide table");
            PG_BTRY(b, hipMemset(b->grp.cnt, 0, (ns_ + 1) * 4), "clear grouping");
            b->grp_cap = cap;
            b->grp_scenes = b->guides_cap = ns_;
        }
        std::vector<MlpView> gv(ns_);
        for (uint32_t i = 0; i < ns_; ++i) {
            const SceneGuide* g = b->scenes[i]->guide.get();
            if (g) gv[i] = g->view;
            else std::memset(&gv[i], 0, sizeof(MlpView));
        }
        PG_BTRY(b, hipMemcpyAsync(b->d_guides, gv.data(), ns_ * sizeof(MlpView), hipMemcpyHostToDevice, s),
                "guide table");
        PG_BTRY(b, hipStreamSynchronize(s), "guide table");
    }
    // passes of ns samples (the captured graph), then one pass of the samples left over;
    // a render of fewer samples than a pass is one whole pass of its own size
    uint32_t ns = b->samples_per_pass();
    if (n_spp > 0 && n_spp < ns) ns = n_spp;
    const uint32_t full = n_spp / ns, rest = n_spp % ns;
    if (b->cfg.use_graph && full > 0 && (!b->graph_ok || !same_params(prm, b->gparams) || b->graph_ns != ns)) {
        b->drop_graph();
        PG_BTRY(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        hipError_t e = enqueue_pass(b, prm, ns);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return b->hip(e, "capture pass");
        }
        PG_BTRY(b, e2, "end capture");
        b->graph = g;
        PG_BTRY(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "instantiate graph");
        (void)hipGraphGetNodes(g, nullptr, &b->graph_nodes);
        b->graph_ok = true;
        b->gparams = prm;
        b->graph_ns = ns;
    }
    PG_BTRY(b, hipEventRecord(b->e0, s), "event");
    for (uint32_t k = 0; k < full; ++k) {
        if (b->cfg.use_graph) PG_BTRY(b, hipGraphLaunch(b->exec, s), "graph launch");
        else PG_BTRY(b, enqueue_pass(b, prm, ns), "batch pass");
    }
    if (rest) PG_BTRY(b, enqueue_pass(b, prm, rest), "batch pass");
    PG_BTRY(b, hipEventRecord(b->e1, s), "event");
    PG_BTRY(b, hipStreamSynchronize(s), "batch render");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->e0, b->e1);
    b->gpu_ms += ms;
    b->spp += n_spp;
    b->paths += (uint64_t)b->total * n_spp;
    return PG_OK;
}

pg_status clear(pg_batch* b) {
    pg_status st = prepare(b);
    if (st != PG_OK) return st;
    PG_BTRY(b, launch_clear_film(b->film, b->total, b->stream()), "clear film");
    PG_BTRY(b, hipMemsetAsync(b->d_sample, 0, 4, b->stream()), "reset sample");
    PG_BTRY(b, hipStreamSynchronize(b->stream()), "sync");
    b->spp = 0;
    return PG_OK;
}

}  // namespace

extern "C" {

pg_status pg_batch_create(pg_ctx* ctx, pg_batch** out) {
    if (!ctx || !out) return PG_ERR_INVALID;
    *out = nullptr;
    if (hipSetDevice(pg_internal::device(ctx)) != hipSuccess) return pg_internal::fail(ctx, PG_ERR_DEVICE, "hipSetDevice");
    std::unique_ptr<pg_batch> b(new (std::nothrow) pg_batch());
    if (!b) return PG_ERR_NOMEM;
    b->ctx = ctx;
    PG_BTRY(b.get(), balloc(&b->d_counters, 8), "alloc batch counters");
    PG_BTRY(b.get(), balloc(&b->d_stats, 8), "alloc batch stats");
    PG_BTRY(b.get(), balloc(&b->d_sample, 1), "alloc batch sample");
    PG_BTRY(b.get(), hipMemset(b->d_counters, 0, 32), "init counters");
    PG_BTRY(b.get(), hipMemset(b->d_stats, 0, 64), "init stats");
    PG_BTRY(b.get(), hipMemset(b->d_sample, 0, 4), "init sample");
    PG_BTRY(b.get(), hipEventCreate(&b->e0), "event");
    PG_BTRY(b.get(), hipEventCreate(&b->e1), "event");
    *out = b.release();
    return PG_OK;
}

pg_status pg_batch_destroy(pg_batch* b) {
    if (!b) return PG_OK;
    (void)hipSetDevice(pg_internal::device(b->ctx));
    (void)hipStreamSynchronize(b->stream());
    b->drop_graph();
    b->free_paths();
    bfree(b->d_scenes); bfree(b->d_views); bfree(b->film); bfree(b->mix);
    bfree(b->grp.cnt); bfree(b->grp.off); bfree(b->grp.cur); bfree(b->grp.perm); bfree(b->grp.n_tiles);
    bfree(b->d_guides);
    bfree(b->d_counters); bfree(b->d_stats); bfree(b->d_sample);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    delete b;
    return PG_OK;
}

pg_status pg_batch_add_scene(pg_batch* b, const float* positions, const float* normals, const uint32_t* material_index,
                             uint32_t n_tris, const pg_material* materials, uint32_t n_materials, const float* env_rgb,
                             int32_t env_width, int32_t env_height, uint32_t* scene_id) {
    if (!b) return PG_ERR_INVALID;
    if (hipSetDevice(pg_internal::device(b->ctx)) != hipSuccess) return b->fail(PG_ERR_DEVICE, "hipSetDevice");
    std::unique_ptr<SceneBuf> s(new (std::nothrow) SceneBuf());
    if (!s) return PG_ERR_NOMEM;
    pg_camera_desc cam{{0.f, 0.f, 1.f}, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, 45.f, 1.f, 1.f};
    std::string e = prepare_scene(positions, normals, material_index, n_tris, materials, n_materials, env_rgb,
                                  env_width, env_height, cam, s->host);
    if (!e.empty()) return b->fail(PG_ERR_INVALID, e);
    if (s->host.bvh.max_depth > 64)
        return b->fail(PG_ERR_OVERFLOW, "BVH deeper than the 64-entry traversal stack (Geometry.h:461)");
    PG_BTRY(b, s->upload(), "upload batch scene");
    b->max_depth = std::max(b->max_depth, (int)s->host.bvh.max_depth);
    if (scene_id) *scene_id = (uint32_t)b->scenes.size();
    b->scenes.push_back(std::move(s));
    b->dirty = true;
    return PG_OK;
}

pg_status pg_batch_scene_textures(pg_batch* b, uint32_t scene_id, const float* uv, const int32_t* material_texture,
                                  uint32_t n_materials, const float* texels, const uint64_t* texel_offset,
                                  const int32_t* tex_width, const int32_t* tex_height, uint32_t n_textures) {
    if (!b) return PG_ERR_INVALID;
    if (scene_id >= b->scenes.size()) return b->fail(PG_ERR_INVALID, "unknown scene");
    if (hipSetDevice(pg_internal::device(b->ctx)) != hipSuccess) return b->fail(PG_ERR_DEVICE, "hipSetDevice");
    PG_BTRY(b, hipStreamSynchronize(pg_internal::stream(b->ctx)), "sync");
    SceneBuf& s = *b->scenes[scene_id];
    pg_status code = PG_OK;
    std::string e = pg_internal::upload_textures(s.tex, s.host.shade.size(), s.host.mats.size(), uv, material_texture,
                                                 n_materials, texels, texel_offset, tex_width, tex_height, n_textures,
                                                 &code);
    if (!e.empty()) {
        s.dev.tex_texels = nullptr; s.dev.tex_info = nullptr; s.dev.uv = nullptr;
        b->dirty = true;
        return b->fail(code, e);
    }
    for (uint32_t i = 0; i < n_materials; ++i) s.host.mats[i].tex = material_texture[i];
    PG_BTRY(b, hipMemcpy(s.mats, s.host.mats.data(), s.host.mats.size() * sizeof(MaterialRec), hipMemcpyHostToDevice),
            "copy mats");
    s.dev.tex_texels = s.tex.texels; s.dev.tex_info = s.tex.info; s.dev.uv = s.tex.uv;
    b->dirty = true;
    return PG_OK;
}

pg_status pg_batch_add_view(pg_batch* b, uint32_t scene_id, const pg_camera_desc* camera, uint32_t* view_id) {
    if (!b || !camera) return PG_ERR_INVALID;
    if (scene_id >= b->scenes.size()) return b->fail(PG_ERR_INVALID, "unknown scene");
    uint32_t w = (uint32_t)camera->width, h = (uint32_t)camera->height;
    if (w == 0 || h == 0 || (float)w != camera->width || (float)h != camera->height)
        return b->fail(PG_ERR_INVALID, "view resolution must be positive integers");
    if ((uint64_t)b->total + (uint64_t)w * h > 0x7fffffffull) return b->fail(PG_ERR_OVERFLOW, "more than 2^31 pixels");
    BatchView v;
    v.cam = make_camera(*camera);
    v.scene = scene_id;
    v.width = w;
    v.height = h;
    v.pixel_offset = b->total;
    if (view_id) *view_id = (uint32_t)b->views.size();
    b->views.push_back(v);
    b->total += w * h;
    b->dirty = true;
    b->film_cap = 0;  // the film layout changes
    return PG_OK;
}

pg_status pg_batch_configure(pg_batch* b, const pg_batch_config* cfg) {
    if (!b) return PG_ERR_INVALID;
    pg_batch_config c = cfg ? *cfg : pg_batch_config{0, 0, 0, 0, 1, 0, 0, 0};
    if (c.stack && (int)c.stack < b->max_depth) return b->fail(PG_ERR_INVALID, "stack shallower than the deepest BVH");
    if (c.stack > 64) return b->fail(PG_ERR_INVALID, "stack above 64 entries");
    if (c.trace_blocks_per_cu > 32 || c.shade_blocks_per_cu > 32) return b->fail(PG_ERR_INVALID, "blocks per CU");
    if (c.samples_per_pass > 64) return b->fail(PG_ERR_INVALID, "samples_per_pass above 64");
    if (c.stack_lds % 16 != 0 || c.stack_lds > 64) return b->fail(PG_ERR_INVALID, "stack_lds must be 0, 16, 32, 48 or 64");
    b->cfg = c;
    b->dirty = true;
    return PG_OK;
}

pg_status pg_batch_clear(pg_batch* b) {
    if (!b) return PG_ERR_INVALID;
    return clear(b);
}

pg_status pg_batch_render(pg_batch* b, const pg_render_params* params, uint32_t n_spp) {
    if (!b) return PG_ERR_INVALID;
    pg_render_params prm = params ? *params : pg_render_params_default();
    pg_status st = check_params(b, prm);
    if (st != PG_OK) return st;
    return render(b, prm, n_spp);
}

pg_status pg_batch_set_guide(pg_batch* b, uint32_t scene, const pg_mlp_config* cfg, const float* params,
                             const float* grid) {
    if (!b) return PG_ERR_INVALID;
    if (scene >= b->scenes.size()) return b->fail(PG_ERR_INVALID, "unknown scene");
    if (!cfg || !params || !grid) return b->fail(PG_ERR_INVALID, "null guide configuration or parameters");
    if (cfg->log2_table < 12 || cfg->log2_table > 22) return b->fail(PG_ERR_INVALID, "log2_table outside 12..22");
    if (!(cfg->res_min >= 1.f) || !(cfg->res_max >= cfg->res_min) || !(cfg->res_max <= 1048576.f))
        return b->fail(PG_ERR_INVALID, "grid resolutions");
    if (hipSetDevice(pg_internal::device(b->ctx)) != hipSuccess) return b->fail(PG_ERR_DEVICE, "hipSetDevice");
    hipStream_t s = b->stream();
    PG_BTRY(b, hipStreamSynchronize(s), "sync");
    b->drop_graph();
    std::unique_ptr<SceneGuide> g(new (std::nothrow) SceneGuide());
    if (!g) return PG_ERR_NOMEM;
    const size_t n_grid = (size_t)kHashLevels * ((size_t)1 << cfg->log2_table) * 2;
    PG_BTRY(b, balloc(&g->params, kParams), "alloc guide");
    PG_BTRY(b, balloc(&g->mirror, kMirror), "alloc guide");
    PG_BTRY(b, balloc(&g->grid, n_grid), "alloc guide");
    PG_BTRY(b, balloc(&g->grid_bf, n_grid / 2), "alloc guide");
    PG_BTRY(b, hipMemcpy(g->params, params, kParams * 4, hipMemcpyHostToDevice), "upload guide");
    PG_BTRY(b, hipMemcpy(g->grid, grid, n_grid * 4, hipMemcpyHostToDevice), "upload guide");
    PG_BTRY(b, launch_mlp_mirror(g->params, g->mirror, s), "guide mirror");
    PG_BTRY(b, launch_grid_mirror(g->grid, g->grid_bf, n_grid, s), "guide mirror");
    PG_BTRY(b, hipStreamSynchronize(s), "guide upload");
    MlpView& v = g->view;
    v.wmirror = g->mirror;
    v.params = g->params;
    v.grid = g->grid_bf;
    const PreparedScene& h = b->scenes[scene]->host;
    mlp_view_frame(v, cfg->log2_table, cfg->res_min, cfg->res_max, h.centre, h.radius);
    b->scenes[scene]->guide = std::move(g);
    return PG_OK;
}

pg_status pg_batch_film_read(pg_batch* b, uint32_t view, float* rgb_sum, uint32_t* spp) {
    if (!b) return PG_ERR_INVALID;
    if (view >= b->views.size()) return b->fail(PG_ERR_INVALID, "unknown view");
    pg_status st = prepare(b);
    if (st != PG_OK) return st;
    const BatchView& v = b->views[view];
    if (rgb_sum) {
        std::vector<float4> tmp((size_t)v.width * v.height);
        PG_BTRY(b, hipMemcpy(tmp.data(), b->film + v.pixel_offset, tmp.size() * 16, hipMemcpyDeviceToHost), "read film");
        for (size_t i = 0; i < tmp.size(); ++i) {
            rgb_sum[3 * i] = tmp[i].x;
            rgb_sum[3 * i + 1] = tmp[i].y;
            rgb_sum[3 * i + 2] = tmp[i].z;
        }
    }
    if (spp) *spp = b->spp;
    return PG_OK;
}

pg_status pg_batch_stats_read(pg_batch* b, pg_batch_stats* s) {
    if (!b || !s) return PG_ERR_INVALID;
    unsigned long long h[8];
    PG_BTRY(b, hipMemcpy(h, b->d_stats, sizeof(h), hipMemcpyDeviceToHost), "read stats");
    std::memset(s, 0, sizeof(*s));
    s->paths = b->paths;
    s->closest_rays = h[0];
    s->shadow_rays = h[1];
    s->gpu_ms = b->gpu_ms;
    s->views = (uint32_t)b->views.size();
    s->scenes = (uint32_t)b->scenes.size();
    s->chunks = (uint32_t)b->chunks.size();
    s->graph_nodes = (uint32_t)b->graph_nodes;
    return PG_OK;
}

pg_status pg_batch_autotune(pg_batch* b, const pg_render_params* params, pg_batch_config* best_out, double* ms_out) {
    if (!b) return PG_ERR_INVALID;
    pg_render_params prm = params ? *params : pg_render_params_default();
    pg_status cst = check_params(b, prm);
    if (cst != PG_OK) return cst;
    uint32_t largest = 0;
    for (const BatchView& v : b->views) largest = std::max(largest, v.width * v.height);
    const int base_stack = stack_entries(std::max(b->max_depth, 1));
    // GPU milliseconds per SPP of c over one pass of min(samples per pass, cap) samples
    // (at least 8 samples in whole passes), after a one-sample warm-up pass that
    // allocates the path state
    auto measure = [&](const pg_batch_config& c, uint32_t cap, double& ms) -> pg_status {
        pg_status st = pg_batch_configure(b, &c);
        if (st != PG_OK) return st;
        if ((st = clear(b)) != PG_OK) return st;
        if ((st = render(b, prm, 1)) != PG_OK) return st;
        const uint32_t k = std::min(c.samples_per_pass ? c.samples_per_pass : 1u, cap);
        const uint32_t timed = k >= 8u ? k : 8u / k * k;
        double before = b->gpu_ms;
        if ((st = render(b, prm, timed)) != PG_OK) return st;
        ms = (b->gpu_ms - before) / (double)timed;
        return PG_OK;
    };
    // paths of the largest chunk under queue size q (the greedy packing of prepare)
    auto chunk_paths = [&](uint32_t q_paths) -> uint64_t {
        const uint64_t q = q_paths ? q_paths : ~0ull;
        uint64_t n = 0, big = 0;
        for (const BatchView& v : b->views) {
            const uint64_t px = (uint64_t)v.width * v.height;
            if (n > 0 && n + px > q) {
                big = std::max(big, n);
                n = 0;
            }
            n += px;
        }
        return std::max(big, n);
    };
    // the most samples per pass (a power of two up to 64) whose chunks stay within the
    // 2^26-path state
    auto most_spp = [&](uint32_t q) -> uint32_t {
        uint32_t spp = 64;
        while (spp > 1 && chunk_paths(q) * spp > (1ull << 26)) spp >>= 1;
        return spp;
    };
    // Small chunks keep one scene's BVH and triangles (at most ~1M triangles, ~75 MB)
    // resident in L2 and the Infinity Cache while its views are traced, and many samples
    // per pass keep each chunk's deep bounces full. Start there, with the single-view
    // tuner's usual pick for the traversal (persistent lanes, 12 blocks per CU, 16 stack
    // entries per lane in LDS), then descend coordinate by coordinate.
    const uint32_t q1 = std::max(largest, 1u << 20), q2 = std::max(largest, 1u << 21);
    pg_batch_config best{q1, 12u, 6u, (uint32_t)base_stack, 1u, 1u, most_spp(q1), 16u};
    // the chunk size and samples per pass are timed over whole passes, the other
    // coordinates over passes of at most 16 samples (their ranking does not depend on
    // the pass length, and the tuner takes a fraction of the time)
    uint32_t cap = 64u;
    double best_ms = 0.0;
    pg_status st = measure(best, cap, best_ms);
    if (st != PG_OK) return st;
    auto consider = [&](pg_batch_config c) -> pg_status {
        if (std::memcmp(&c, &best, sizeof(c)) == 0) return PG_OK;
        double ms = 0.0;
        pg_status s2 = measure(c, cap, ms);
        if (s2 != PG_OK) return s2;
        // graph replay stays on unless eager launches are clearly faster: the graph takes
        // the per-launch host work off the render (hundreds of launches per pass)
        const double margin = (best.use_graph && !c.use_graph) ? 0.98 : 1.0;
        if (ms < best_ms * margin) {
            best_ms = ms;
            best = c;
        }
        return PG_OK;
    };
    // chunk size with its most samples per pass, then fewer samples per pass
    for (uint32_t q : {q2, 0u}) {
        pg_batch_config c = best;
        c.queue_paths = q;
        c.samples_per_pass = most_spp(q);
        if ((st = consider(c)) != PG_OK) return st;
    }
    for (uint32_t div : {2u, 4u}) {
        pg_batch_config c = best;
        if (most_spp(c.queue_paths) / div < 1u) continue;
        c.samples_per_pass = most_spp(c.queue_paths) / div;
        if ((st = consider(c)) != PG_OK) return st;
    }
    cap = 16u;
    if ((st = measure(best, cap, best_ms)) != PG_OK) return st;
    std::vector<std::pair<uint32_t pg_batch_config::*, std::vector<uint32_t>>> cand = {
        {&pg_batch_config::use_graph, {0u, 1u}},
        {&pg_batch_config::dyn_fetch, {0u, 1u}},
        {&pg_batch_config::trace_blocks_per_cu, {8u, 16u}},
        {&pg_batch_config::shade_blocks_per_cu, {3u, 4u}},
        {&pg_batch_config::stack_lds, {32u}},
    };
    for (auto& kv : cand) {
        for (uint32_t val : kv.second) {
            pg_batch_config c = best;
            c.*(kv.first) = val;
            if ((st = consider(c)) != PG_OK) return st;
        }
    }
    // the pick's time over one whole pass
    if ((st = measure(best, 64u, best_ms)) != PG_OK) return st;
    if ((st = pg_batch_configure(b, &best)) != PG_OK) return st;
    if ((st = clear(b)) != PG_OK) return st;
    b->gpu_ms = 0.0;
    b->paths = 0;
    PG_BTRY(b, hipMemset(b->d_stats, 0, 64), "reset stats");
    if (best_out) *best_out = best;
    if (ms_out) *ms_out = best_ms;
    return PG_OK;
}

}  // extern "C"
