// api.hip — C ABI (include/pg_tracer.h): device context, scene upload, the
// wavefront render loop (RayTracer::render drop-in) and test queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pg_tracer.h"
#include "host_prep.hpp"
#include "kernels.h"
#include "guiding.h"
#include "mlp.h"
#include "ctx_internal.h"
#include "bvh_build.h"

using namespace pg;

static __device__ __align__(256) float scratch[(size_t)(1u << 18) * 16];

namespace {
thread_local std::string g_last_error;

// Staging for pg_trace_closest_noalloc: per ray the ray (6 floats), the closest-hit
// arrays the traversal kernel writes (id, t, alpha, beta, gamma) and the packed
// pg_hit records (5 words): 16 words, 64 B. One array per GPU, held until the
// process ends; its device address comes from hipGetSymbolAddress (the host-side
// name is not a device address).
constexpr uint32_t kNoAllocMaxRays = 1u << 18;
static_assert(sizeof(scratch) == (size_t)kNoAllocMaxRays * 16 * sizeof(float), "64 B of scratch per ray");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch[kMaxDevices] = {};

// Device address of this GPU's scratch array (the device is current); nullptr on error.
float* scratch_for_device(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[device] = (float*)ptr;
    }
    return g_scratch[device];
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}
template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

struct TimedLaunch {
    int kind;  // 0 extend, 1 shade, 2 shadow, 3 other
    hipEvent_t a, b;
};
}  // namespace

struct pg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t bvh_builder = PG_BVH_HOST;  // pg_set_bvh_builder
    float* noalloc_scratch = nullptr;  // this device's scratch array (pg_trace_closest_noalloc)
    std::string err;
    int cu_count = 0;
    std::string name;
    // scene
    PreparedScene host;
    bool has_scene = false;
    BvhNode* d_nodes = nullptr;
    TriIsect* d_tris = nullptr;
    TriShade* d_shade = nullptr;
    MaterialRec* d_mats = nullptr;
    LightRec* d_lights = nullptr;
    float4* d_env_tex = nullptr;
    float* d_env_pdf = nullptr;
    float* d_env_prob = nullptr;
    uint32_t* d_env_alias = nullptr;
    pg_internal::TextureSet tex;
    uint32_t* leaf_nbrs = nullptr;
    uint32_t leaf_nbrs_cap = 0;
    DevScene dev{};
    // path state + film
    PathState ps{};
    uint32_t ps_capacity = 0;
    float4* film = nullptr;
    uint32_t film_n = 0;
    uint32_t spp = 0;
    unsigned long long* d_stats = nullptr;
    uint32_t* d_counters = nullptr;
    bool counting = false;
    bool timing = true;
    // host-side stats
    double ms[5] = {0, 0, 0, 0, 0};
    double ms_total = 0;
    uint64_t n_extend = 0, n_shadow = 0, paths = 0;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    std::vector<TimedLaunch> launches;
    // record pass (forward vertices per path, back-propagated into records)
    RecordState rs{};
    uint32_t rs_depth = 0, rs_paths = 0;
    uint32_t* rec_counts = nullptr;
    uint32_t* rec_offsets = nullptr;
    uint32_t* scan_tmp = nullptr;
    uint32_t* d_total = nullptr;
    uint32_t scan_cap = 0;
    Record* records = nullptr;
    uint64_t n_records = 0, rec_capacity = 0;
    // radiance cache
    CacheBuffers cb{};
    uint32_t cache_cap = 0, cache_pow2_cap = 0, cache_cells_cap = 0, cache_stride = 0;
    GuideView gv{};
    bool cache_ready = false;
    double cache_build_ms = 0.0;
    uint32_t* guide_cells = nullptr;
    uint32_t guide_cells_cap = 0;
    // kNN query order per bounce: Morton keys of the shading points, radix-sorted
    uint32_t *q_keys = nullptr, *q_vals = nullptr, *q_keys_tmp = nullptr, *q_vals_tmp = nullptr;
    uint32_t *q_hist = nullptr, *q_scan = nullptr;
    uint32_t q_cap = 0;
    // training data beside the records (records [0, n_aux) carry it)
    TrainAux* rec_aux = nullptr;
    uint64_t n_aux = 0;
    uint64_t rec_gen = 0;  // bumped whenever the record set changes
    // learned guide
    struct Mlp {
        bool ready = false, diverged = false;
        pg_mlp_config cfg{};
        MlpView view{};
        float* params = nullptr;
        float* pm = nullptr;
        float* pv = nullptr;
        uint16_t* mirror = nullptr;
        float* grid = nullptr;
        float* gm = nullptr;
        float* gv = nullptr;
        float* ggrad = nullptr;
        uint32_t* grid_bf = nullptr;
        size_t n_grid = 0;
        size_t n_ggrad = 0;  // floats of the privatised gradient accumulators
        float* partial = nullptr;
        uint32_t groups_cap = 0;
        TrainSample* batch = nullptr;
        TrainSample* batch_raw = nullptr;  // gathered samples before the spatial sort
        SortBuffers sort{};
        uint32_t batch_cap = 0;
        uint16_t* rows = nullptr;
        uint32_t rows_cap = 0;
        float* loss = nullptr;
        uint32_t loss_cap = 0;
        float* scal = nullptr;
        uint32_t scal_cap = 0;
        uint32_t* status = nullptr;
        double* sum_part = nullptr;
        uint32_t step = 0, chunks = 1, fallbacks = 0;
        float last_loss = 0.f, w_scale = 1.f, train_ms = 0.f;
        // rising-loss detector: smoothed loss, its best value, steps above best + rise
        float loss_ema = 0.f, loss_best = 0.f;
        uint32_t loss_seen = 0, rise_run = 0, diverge_reason = 0;
        uint64_t w_scale_gen = ~0ull;
        uint32_t window_passes = 0, online_passes = 0, record_passes = 0;
    } mlp;
    float4* mix = nullptr;
    uint32_t mix_cap = 0;

    pg_status fail(pg_status code, const std::string& m) {
        err = m;
        g_last_error = m;
        return code;
    }
    pg_status hip(hipError_t e, const char* what) {
        if (e == hipSuccess) return PG_OK;
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_DEVICE,
                    std::string(what) + ": " + hipGetErrorString(e));
    }
    hipEvent_t ev() {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[pool_used++];
    }
    void free_scene() {
        dfree(d_nodes); dfree(d_tris); dfree(d_shade); dfree(d_mats); dfree(d_lights);
        dfree(d_env_tex); dfree(d_env_pdf); dfree(d_env_prob); dfree(d_env_alias);
        tex.release();
        dev.tex_texels = nullptr; dev.tex_info = nullptr; dev.uv = nullptr;
        has_scene = false;
    }
    void free_records() {
        dfree(rs.vpos); dfree(rs.vnrm); dfree(rs.vwi); dfree(rs.vwo); dfree(rs.vw); dfree(rs.ve); dfree(rs.vme);
        dfree(rs.vd);
        dfree(rs.nv);
        rs_depth = rs_paths = 0;
        dfree(rec_counts); dfree(rec_offsets); dfree(scan_tmp); dfree(d_total);
        scan_cap = 0;
    }
    void free_cache() {
        dfree(cb.keys); dfree(cb.vals); dfree(cb.keys_tmp); dfree(cb.vals_tmp); dfree(cb.hist); dfree(cb.scan_tmp);
        dfree(cb.pos); dfree(cb.dir); dfree(cb.nodes); dfree(cb.qt); dfree(cb.qt_acc);
        cache_cap = cache_pow2_cap = cache_cells_cap = cache_stride = 0;
        dfree(leaf_nbrs);
        leaf_nbrs_cap = 0;
        cache_ready = false;
        gv = GuideView{};
    }
    void free_mlp() {
        Mlp& m = mlp;
        dfree(m.params); dfree(m.pm); dfree(m.pv); dfree(m.mirror);
        dfree(m.grid); dfree(m.gm); dfree(m.gv); dfree(m.ggrad); dfree(m.grid_bf);
        dfree(m.partial); dfree(m.batch); dfree(m.rows); dfree(m.loss); dfree(m.scal); dfree(m.status); dfree(m.sum_part);
        dfree(m.batch_raw); dfree(m.sort.keys); dfree(m.sort.vals); dfree(m.sort.keys_tmp); dfree(m.sort.vals_tmp);
        dfree(m.sort.hist); dfree(m.sort.scan);
        m = Mlp{};
    }
    void free_paths() {
        dfree(ps.beta); dfree(ps.radiance); dfree(ps.film_pos); dfree(ps.info);
        for (int k = 0; k < 2; ++k) { dfree(ps.queue[k]); dfree(ps.ray_o[k]); dfree(ps.ray_d[k]); }
        dfree(ps.hit); dfree(ps.sh_o); dfree(ps.sh_d); dfree(ps.sh_c); dfree(ps.spill);
        ps.spill_lanes = 0;
        ps_capacity = 0;
    }
};

namespace pg_internal {
hipStream_t stream(pg_ctx* c) { return c->stream; }
int cu_count(pg_ctx* c) { return c->cu_count; }
int device(pg_ctx* c) { return c->device; }
pg_status fail(pg_ctx* c, pg_status code, const std::string& m) { return c->fail(code, m); }

void TextureSet::release() {
    if (texels) (void)hipFree(texels);
    if (info) (void)hipFree(info);
    if (uv) (void)hipFree(uv);
    texels = nullptr; info = nullptr; uv = nullptr;
}

std::string upload_textures(TextureSet& set, size_t n_tris, size_t n_mats, const float* uv,
                            const int32_t* material_texture, uint32_t n_materials, const float* texels,
                            const uint64_t* texel_offset, const int32_t* tex_width, const int32_t* tex_height,
                            uint32_t n_textures, pg_status* code) {
    *code = PG_ERR_INVALID;
    if (n_materials != n_mats) return "material count differs from the uploaded scene";
    if (!material_texture || !uv) return "null uv or material_texture";
    if (n_textures > 0 && (!texels || !texel_offset || !tex_width || !tex_height)) return "null texture array";
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_textures; ++i) {
        if (tex_width[i] < 1 || tex_height[i] < 1) return "texture size must be >= 1x1";
        uint64_t n = (uint64_t)tex_width[i] * (uint64_t)tex_height[i];
        if (texel_offset[i] + n > total) total = texel_offset[i] + n;
    }
    if (total >= (1ull << 31)) return "more than 2^31 texels";
    for (uint32_t i = 0; i < n_materials; ++i)
        if (material_texture[i] < -1 || material_texture[i] >= (int32_t)n_textures)
            return "material texture index out of range";
    set.release();
    std::vector<float4> tx((size_t)total);
    for (uint64_t k = 0; k < total; ++k) tx[k] = make_float4(texels[3 * k], texels[3 * k + 1], texels[3 * k + 2], 0.f);
    std::vector<int4> info(n_textures ? n_textures : 1, make_int4(0, 1, 1, 0));
    for (uint32_t i = 0; i < n_textures; ++i) info[i] = make_int4((int)texel_offset[i], tex_width[i], tex_height[i], 0);
    *code = PG_ERR_NOMEM;
    if (hipMalloc((void**)&set.texels, (total ? (size_t)total : 1) * sizeof(float4)) != hipSuccess ||
        hipMalloc((void**)&set.info, info.size() * sizeof(int4)) != hipSuccess ||
        hipMalloc((void**)&set.uv, (n_tris ? 3 * n_tris : 1) * sizeof(float2)) != hipSuccess) {
        set.release();
        return "alloc textures";
    }
    *code = PG_ERR_DEVICE;
    if ((total && hipMemcpy(set.texels, tx.data(), tx.size() * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(set.info, info.data(), info.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess ||
        (n_tris && hipMemcpy(set.uv, uv, n_tris * 6 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)) {
        set.release();
        return "copy textures";
    }
    *code = PG_OK;
    return std::string();
}
}  // namespace pg_internal

#define PG_CHECK_CTX(c)                                  \
    do {                                                 \
        if (!(c)) {                                      \
            g_last_error = "null context";               \
            return PG_ERR_INVALID;                       \
        }                                                \
        if (hipSetDevice((c)->device) != hipSuccess)     \
            return (c)->fail(PG_ERR_DEVICE, "hipSetDevice failed"); \
    } while (0)

#define PG_TRY(c, expr, what)                          \
    do {                                               \
        pg_status _s = (c)->hip((expr), (what));       \
        if (_s != PG_OK) return _s;                    \
    } while (0)

extern "C" {

const char* pg_last_error(const pg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }

pg_status pg_create(pg_ctx** out, int device) {
    if (!out) { g_last_error = "null output"; return PG_ERR_INVALID; }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        g_last_error = "no HIP device available";
        return PG_ERR_DEVICE;
    }
    if (device < 0 || device >= n) { g_last_error = "device index out of range"; return PG_ERR_INVALID; }
    pg_ctx* c = new (std::nothrow) pg_ctx();
    if (!c) return PG_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; g_last_error = "hipSetDevice failed"; return PG_ERR_DEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; g_last_error = "hipGetDeviceProperties failed"; return PG_ERR_DEVICE; }
    c->cu_count = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c; g_last_error = "hipStreamCreate failed"; return PG_ERR_DEVICE;
    }
    if (dalloc(&c->d_stats, 8) != hipSuccess || dalloc(&c->d_counters, 8) != hipSuccess) {
        pg_destroy(c); g_last_error = "device allocation failed"; return PG_ERR_NOMEM;
    }
    c->noalloc_scratch = scratch_for_device(device);
    if (!c->noalloc_scratch) {
        pg_destroy(c); g_last_error = "hipGetSymbolAddress of the scratch array failed"; return PG_ERR_DEVICE;
    }
    (void)hipMemset(c->d_stats, 0, 8 * sizeof(unsigned long long));
    (void)hipMemset(c->d_counters, 0, 8 * sizeof(uint32_t));
    *out = c;
    return PG_OK;
}

pg_status pg_destroy(pg_ctx* c) {
    if (!c) return PG_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->free_scene();
    c->free_paths();
    c->free_records();
    c->free_cache();
    dfree(c->records);
    dfree(c->rec_aux);
    c->free_mlp();
    dfree(c->mix);
    dfree(c->guide_cells);
    dfree(c->q_keys); dfree(c->q_vals); dfree(c->q_keys_tmp); dfree(c->q_vals_tmp); dfree(c->q_hist); dfree(c->q_scan);
    dfree(c->film);
    dfree(c->d_stats);
    dfree(c->d_counters);
    for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PG_OK;
}

pg_status pg_device_info(pg_ctx* c, char* name, size_t name_len, int32_t* cu_count, double* peak_hbm_gbs) {
    PG_CHECK_CTX(c);
    if (name && name_len) {
        std::strncpy(name, c->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = c->cu_count;
    if (peak_hbm_gbs) *peak_hbm_gbs = 8000.0;  // MI355X HBM3E published peak
    return PG_OK;
}

// The GPU binned-SAH build of n triangles' positions (host array): the same arrays
// as the host build, in the reference's topology; build_ms holds the GPU time.
static pg_status build_bvh_on_device(pg_ctx* c, const float* positions, uint32_t n, HostBvh& out) {
    float* dpos = nullptr;
    pg_status st = c->hip(dalloc(&dpos, (size_t)n * 9), "alloc positions");
    if (st == PG_OK)
        st = c->hip(hipMemcpy(dpos, positions, (size_t)n * 36, hipMemcpyHostToDevice), "copy positions");
    double ms = 0.0;
    if (st == PG_OK) {
        try {
            st = c->hip(gpu_build_bvh(dpos, n, out, &ms, c->stream), "GPU BVH build");
        } catch (const std::bad_alloc&) {
            st = c->fail(PG_ERR_NOMEM, "out of host memory for the BVH");
        }
    }
    dfree(dpos);
    out.build_ms = ms;
    return st;
}

pg_status pg_scene_upload(pg_ctx* c, const float* positions, const float* normals, const uint32_t* material_index,
                          uint32_t n_tris, const pg_material* materials, uint32_t n_materials, const float* env_rgb,
                          int32_t env_width, int32_t env_height, const pg_camera_desc* camera) {
    PG_CHECK_CTX(c);
    if (!camera) return c->fail(PG_ERR_INVALID, "null camera");
    PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
    c->free_scene();
    c->host = PreparedScene();
    std::string e;
    HostBvh gpu_bvh;
    const bool on_gpu = c->bvh_builder == PG_BVH_GPU && positions && n_tris > 0 && n_tris <= (1u << 26);
    if (on_gpu) {
        pg_status st = build_bvh_on_device(c, positions, n_tris, gpu_bvh);
        if (st != PG_OK) return st;
    }
    try {
        e = prepare_scene(positions, normals, material_index, n_tris, materials, n_materials, env_rgb, env_width,
                          env_height, *camera, c->host, on_gpu ? &gpu_bvh : nullptr);
    } catch (const std::bad_alloc&) {
        c->host = PreparedScene();
        return c->fail(PG_ERR_NOMEM, "out of host memory preparing the scene");
    } catch (const std::exception& x) {
        c->host = PreparedScene();
        return c->fail(PG_ERR_INVALID, std::string("scene preparation failed: ") + x.what());
    }
    if (!e.empty()) return c->fail(PG_ERR_INVALID, e);
    if (c->host.bvh.max_depth > 64)
        return c->fail(PG_ERR_OVERFLOW, "BVH deeper than the 64-entry traversal stack (Geometry.h:461)");
    PreparedScene& h = c->host;
    PG_TRY(c, dalloc(&c->d_nodes, h.node_slots.size()), "alloc nodes");
    // one zeroed record past the end: traversal reads 64 bytes at every triangle
    PG_TRY(c, dalloc(&c->d_tris, h.tris.size() + 1), "alloc tris");
    PG_TRY(c, hipMemset(c->d_tris + h.tris.size(), 0, sizeof(TriIsect)), "clear tris pad");
    PG_TRY(c, dalloc(&c->d_shade, h.shade.size()), "alloc shade");
    PG_TRY(c, dalloc(&c->d_mats, h.mats.size()), "alloc mats");
    PG_TRY(c, dalloc(&c->d_lights, h.lights.size() ? h.lights.size() : 1), "alloc lights");
    PG_TRY(c, hipMemcpy(c->d_nodes, h.node_slots.data(), h.node_slots.size() * sizeof(BvhNode), hipMemcpyHostToDevice), "copy nodes");
    PG_TRY(c, hipMemcpy(c->d_tris, h.tris.data(), h.tris.size() * sizeof(TriIsect), hipMemcpyHostToDevice), "copy tris");
    PG_TRY(c, hipMemcpy(c->d_shade, h.shade.data(), h.shade.size() * sizeof(TriShade), hipMemcpyHostToDevice), "copy shade");
    PG_TRY(c, hipMemcpy(c->d_mats, h.mats.data(), h.mats.size() * sizeof(MaterialRec), hipMemcpyHostToDevice), "copy mats");
    if (!h.lights.empty())
        PG_TRY(c, hipMemcpy(c->d_lights, h.lights.data(), h.lights.size() * sizeof(LightRec), hipMemcpyHostToDevice), "copy lights");
    if (h.has_env) {
        size_t n = (size_t)h.env.w * h.env.h;
        PG_TRY(c, dalloc(&c->d_env_tex, n), "alloc env");
        PG_TRY(c, dalloc(&c->d_env_pdf, n), "alloc env pdf");
        PG_TRY(c, dalloc(&c->d_env_prob, n), "alloc env prob");
        PG_TRY(c, dalloc(&c->d_env_alias, n), "alloc env alias");
        PG_TRY(c, hipMemcpy(c->d_env_tex, h.env.texels.data(), n * 16, hipMemcpyHostToDevice), "copy env");
        PG_TRY(c, hipMemcpy(c->d_env_pdf, h.env.pdf.data(), n * 4, hipMemcpyHostToDevice), "copy env pdf");
        PG_TRY(c, hipMemcpy(c->d_env_prob, h.env.prob.data(), n * 4, hipMemcpyHostToDevice), "copy env prob");
        PG_TRY(c, hipMemcpy(c->d_env_alias, h.env.alias.data(), n * 4, hipMemcpyHostToDevice), "copy env alias");
    }
    DevScene& d = c->dev;
    d.nodes = c->d_nodes; d.tris = c->d_tris; d.shade = c->d_shade; d.mats = c->d_mats; d.lights = c->d_lights;
    d.env_tex = c->d_env_tex; d.env_pdf = c->d_env_pdf; d.env_prob = c->d_env_prob; d.env_alias = c->d_env_alias;
    d.n_lights = (int32_t)h.lights.size();
    d.env_w = h.has_env ? h.env.w : 0;
    d.env_h = h.has_env ? h.env.h : 0;
    d.has_env = h.has_env ? 1 : 0;
    d.finite_bounds = h.finite_bounds ? 1 : 0;
    d.centre[0] = h.centre[0]; d.centre[1] = h.centre[1]; d.centre[2] = h.centre[2];
    d.radius = h.radius;
    d.cam = h.camera;
    c->has_scene = true;
    c->mlp.ready = false;  // the guide's position frame belongs to the previous scene
    // film for the camera resolution
    uint32_t W = (uint32_t)h.camera.width, H = (uint32_t)h.camera.height;
    uint32_t fn = W * H;
    if (fn != c->film_n) {
        dfree(c->film);
        PG_TRY(c, dalloc(&c->film, fn), "alloc film");
        c->film_n = fn;
    }
    PG_TRY(c, launch_clear_film(c->film, fn, c->stream), "clear film");
    c->spp = 0;
    PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
    return PG_OK;
}

pg_status pg_scene_textures(pg_ctx* c, const float* uv, const int32_t* material_texture, uint32_t n_materials,
                            const float* texels, const uint64_t* texel_offset, const int32_t* tex_width,
                            const int32_t* tex_height, uint32_t n_textures) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "pg_scene_textures before pg_scene_upload");
    PreparedScene& h = c->host;
    PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
    pg_status code = PG_OK;
    std::string e = pg_internal::upload_textures(c->tex, h.shade.size(), h.mats.size(), uv, material_texture,
                                                 n_materials, texels, texel_offset, tex_width, tex_height, n_textures,
                                                 &code);
    if (!e.empty()) {
        c->dev.tex_texels = nullptr; c->dev.tex_info = nullptr; c->dev.uv = nullptr;
        return c->fail(code, e);
    }
    for (uint32_t i = 0; i < n_materials; ++i) h.mats[i].tex = material_texture[i];
    PG_TRY(c, hipMemcpy(c->d_mats, h.mats.data(), h.mats.size() * sizeof(MaterialRec), hipMemcpyHostToDevice), "copy mats");
    c->dev.tex_texels = c->tex.texels; c->dev.tex_info = c->tex.info; c->dev.uv = c->tex.uv;
    return PG_OK;
}

pg_status pg_bvh_info(pg_ctx* c, uint32_t* n_nodes, uint32_t* max_depth, double* build_ms) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (n_nodes) *n_nodes = (uint32_t)c->host.bvh.nodes.size();
    if (max_depth) *max_depth = c->host.bvh.max_depth;
    if (build_ms) *build_ms = c->host.bvh.build_ms;
    return PG_OK;
}

pg_status pg_bvh_export(pg_ctx* c, float* nodes, int32_t* tri_index) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (nodes) std::memcpy(nodes, c->host.bvh.nodes.data(), c->host.bvh.nodes.size() * sizeof(BvhNode));
    if (tri_index) std::memcpy(tri_index, c->host.bvh.tri_index.data(), c->host.bvh.tri_index.size() * 4);
    return PG_OK;
}

pg_status pg_bvh_build_host(const float* positions, uint32_t n_tris, uint32_t* n_nodes, float* nodes_out,
                            int32_t* tri_index_out) {
    if (!positions || n_tris == 0) { g_last_error = "empty triangle set"; return PG_ERR_INVALID; }
    HostBvh b;
    try {
        build_bvh(positions, n_tris, b);
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory building the BVH";
        return PG_ERR_NOMEM;
    } catch (const std::exception& x) {
        g_last_error = std::string("BVH build failed: ") + x.what();
        return PG_ERR_INVALID;
    }
    if (n_nodes) *n_nodes = (uint32_t)b.nodes.size();
    if (nodes_out) std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(BvhNode));
    if (tri_index_out) std::memcpy(tri_index_out, b.tri_index.data(), b.tri_index.size() * 4);
    return PG_OK;
}

pg_status pg_bvh_build_gpu(pg_ctx* c, const float* positions, uint32_t n_tris, uint32_t* n_nodes, float* nodes_out,
                           int32_t* tri_index_out, double* build_ms) {
    PG_CHECK_CTX(c);
    if (!positions || n_tris == 0) return c->fail(PG_ERR_INVALID, "empty triangle set");
    if (n_tris > (1u << 26)) return c->fail(PG_ERR_INVALID, "more than 2^26 triangles");
    HostBvh b;
    pg_status st = build_bvh_on_device(c, positions, n_tris, b);
    if (st != PG_OK) return st;
    if (n_nodes) *n_nodes = (uint32_t)b.nodes.size();
    if (nodes_out) std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(BvhNode));
    if (tri_index_out) std::memcpy(tri_index_out, b.tri_index.data(), b.tri_index.size() * 4);
    if (build_ms) *build_ms = b.build_ms;
    return PG_OK;
}

pg_status pg_set_bvh_builder(pg_ctx* c, int32_t builder) {
    PG_CHECK_CTX(c);
    if (builder != PG_BVH_HOST && builder != PG_BVH_GPU) return c->fail(PG_ERR_INVALID, "unknown BVH builder");
    c->bvh_builder = builder;
    return PG_OK;
}

pg_status pg_trace_closest(pg_ctx* c, const pg_ray* rays, uint32_t n, pg_hit* hits) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (n == 0) return PG_OK;
    if (!rays || !hits) return c->fail(PG_ERR_INVALID, "null array");
    float *dr = nullptr, *dt = nullptr, *da = nullptr, *db = nullptr, *dg = nullptr;
    uint32_t* did = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dr, (size_t)n * 6), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&did, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dt, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&da, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&db, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dg, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dr, rays, (size_t)n * 24, hipMemcpyHostToDevice), "copy rays")) != PG_OK) break;
        if ((st = c->hip(launch_trace_closest(c->dev, dr, n, did, dt, da, db, dg, (int)c->host.bvh.max_depth, c->counting, c->d_stats, c->stream),
                         "trace_closest")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        std::vector<uint32_t> id(n);
        std::vector<float> t(n), a(n), b(n), g(n);
        if ((st = c->hip(hipMemcpy(id.data(), did, n * 4, hipMemcpyDeviceToHost), "copy hits")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(t.data(), dt, n * 4, hipMemcpyDeviceToHost), "copy hits")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(a.data(), da, n * 4, hipMemcpyDeviceToHost), "copy hits")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(b.data(), db, n * 4, hipMemcpyDeviceToHost), "copy hits")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(g.data(), dg, n * 4, hipMemcpyDeviceToHost), "copy hits")) != PG_OK) break;
        for (uint32_t i = 0; i < n; ++i) hits[i] = pg_hit{id[i], t[i], a[i], b[i], g[i]};
    } while (0);
    dfree(dr); dfree(did); dfree(dt); dfree(da); dfree(db); dfree(dg);
    return st;
}

pg_status pg_trace_closest_noalloc(pg_ctx* c, const pg_ray* rays, uint32_t n, pg_hit* hits) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (n == 0) return PG_OK;
    if (!rays || !hits) return c->fail(PG_ERR_INVALID, "null array");
    if (n > kNoAllocMaxRays)
        return c->fail(PG_ERR_INVALID, "pg_trace_closest_noalloc: more than 2^18 rays (its scratch limit); "
                                       "split the batch or call pg_trace_closest");
    float* base = c->noalloc_scratch;
    if (!base) return c->fail(PG_ERR_STATE, "no scratch array on this device");
    float* dr = base;                                  // 6n: rays
    uint32_t* did = (uint32_t*)(base + (size_t)6 * n); // n: ids
    float* dt = base + (size_t)7 * n;                  // 4n: t, alpha, beta, gamma
    float* da = dt + n;
    float* db = da + n;
    float* dg = db + n;
    pg_hit* dh = (pg_hit*)(base + (size_t)11 * n);     // 5n: packed hits
    PG_TRY(c, hipMemcpyAsync(dr, rays, (size_t)n * sizeof(pg_ray), hipMemcpyHostToDevice, c->stream), "copy rays");
    PG_TRY(c, launch_trace_closest(c->dev, dr, n, did, dt, da, db, dg, (int)c->host.bvh.max_depth, c->counting,
                                   c->d_stats, c->stream), "trace_closest");
    PG_TRY(c, launch_pack_hits(did, dt, da, db, dg, n, (uint32_t*)dh, c->stream), "pack hits");
    PG_TRY(c, hipMemcpyAsync(hits, dh, (size_t)n * sizeof(pg_hit), hipMemcpyDeviceToHost, c->stream), "copy hits");
    PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
    return PG_OK;
}

pg_status pg_trace_shadow(pg_ctx* c, const pg_ray* rays, const float* max_t, uint32_t n, uint8_t* visible) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (n == 0) return PG_OK;
    if (!rays || !max_t || !visible) return c->fail(PG_ERR_INVALID, "null array");
    float *dr = nullptr, *dm = nullptr;
    uint8_t* dv = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dr, (size_t)n * 6), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dm, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dv, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dr, rays, (size_t)n * 24, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dm, max_t, (size_t)n * 4, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(launch_trace_shadow(c->dev, dr, dm, n, dv, (int)c->host.bvh.max_depth, c->stream), "trace_shadow")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        st = c->hip(hipMemcpy(visible, dv, n, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(dr); dfree(dm); dfree(dv);
    return st;
}

pg_status pg_generate_rays(pg_ctx* c, const float* pxy, uint32_t n, pg_ray* rays) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (n == 0) return PG_OK;
    float *dp = nullptr, *dr = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dp, (size_t)n * 2), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dr, (size_t)n * 6), "alloc")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dp, pxy, (size_t)n * 8, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(launch_generate_rays(c->dev, dp, n, dr, c->stream), "generate_rays")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        st = c->hip(hipMemcpy(rays, dr, (size_t)n * 24, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(dp); dfree(dr);
    return st;
}

pg_status pg_rng_values(pg_ctx* c, uint32_t seed, uint32_t sample, const uint32_t* pixel, const uint32_t* dim,
                        uint32_t n, float* out) {
    PG_CHECK_CTX(c);
    if (n == 0) return PG_OK;
    uint32_t *dp = nullptr, *dd = nullptr;
    float* dout = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dp, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dd, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dout, n), "alloc")) != PG_OK) break;
        hipMemcpy(dp, pixel, n * 4, hipMemcpyHostToDevice);
        hipMemcpy(dd, dim, n * 4, hipMemcpyHostToDevice);
        if ((st = c->hip(launch_rng(seed, sample, dp, dd, n, dout, c->stream), "rng")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        st = c->hip(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(dp); dfree(dd); dfree(dout);
    return st;
}

pg_status pg_bsdf_query(pg_ctx* c, const pg_material* material, uint32_t n, const float* sd, const float* u,
                        const float* wi_eval, float* out) {
    PG_CHECK_CTX(c);
    if (n == 0) return PG_OK;
    if (!material || !sd || !u || !wi_eval || !out) return c->fail(PG_ERR_INVALID, "null array");
    if (n > (1u << 24)) return c->fail(PG_ERR_INVALID, "more than 2^24 queries");
    MaterialRec m;
    if (!material_record(*material, m)) return c->fail(PG_ERR_INVALID, "unknown material type");
    float *dsd = nullptr, *du = nullptr, *dwe = nullptr, *dout = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dsd, (size_t)n * 12), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&du, (size_t)n * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dwe, (size_t)n * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dout, (size_t)n * 15), "alloc")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dsd, sd, (size_t)n * 48, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(du, u, (size_t)n * 12, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dwe, wi_eval, (size_t)n * 12, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(launch_bsdf_query(m, dsd, du, dwe, n, dout, c->stream), "bsdf query")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        st = c->hip(hipMemcpy(out, dout, (size_t)n * 60, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(dsd); dfree(du); dfree(dwe); dfree(dout);
    return st;
}

pg_render_params pg_render_params_default(void) {
    pg_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.seed = 1;           // MTRandom default seed (Sampling.h:21)
    p.max_depth = 8;      // Renderer.h:568
    p.rr_depth = 3;       // Renderer.h:561
    p.integrator = PG_INTEGRATOR_PATH;
    p.record = 0;
    p.guide_prob = 0.5f;
    p.knn_k = 16;
    p.dyn_fetch = 1;  // launch shape only: images are those of every other shape
    return p;
}

pg_status pg_film_clear(pg_ctx* c) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    PG_TRY(c, launch_clear_film(c->film, c->film_n, c->stream), "clear film");
    PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
    c->spp = 0;
    return PG_OK;
}

static pg_status ensure_paths(pg_ctx* c, uint32_t n) {
    if (c->ps_capacity >= n) return PG_OK;
    c->free_paths();
    PathState& p = c->ps;
    PG_TRY(c, dalloc(&p.beta, n), "alloc paths");
    PG_TRY(c, dalloc(&p.radiance, n), "alloc paths");
    PG_TRY(c, dalloc(&p.film_pos, n), "alloc paths");
    PG_TRY(c, dalloc(&p.info, n), "alloc paths");
    for (int k = 0; k < 2; ++k) {
        PG_TRY(c, dalloc(&p.queue[k], n), "alloc queues");
        PG_TRY(c, dalloc(&p.ray_o[k], n), "alloc queues");
        PG_TRY(c, dalloc(&p.ray_d[k], n), "alloc queues");
    }
    PG_TRY(c, dalloc(&p.hit, n), "alloc hits");
    PG_TRY(c, dalloc(&p.sh_o, n), "alloc shadow");
    PG_TRY(c, dalloc(&p.sh_d, n), "alloc shadow");
    PG_TRY(c, dalloc(&p.sh_c, n), "alloc shadow");
    // traversal stack entries beyond the LDS share, for up to 32 traversal blocks of
    // 128 lanes per CU
    p.spill_lanes = (uint32_t)c->cu_count * 32u * 128u;
    PG_TRY(c, dalloc(&p.spill, (size_t)p.spill_lanes * kSpillEntries), "alloc traversal stacks");
    p.counters = c->d_counters;
    p.stats = c->d_stats;
    p.capacity = n;
    c->ps_capacity = n;
    return PG_OK;
}

static pg_status ensure_scan(pg_ctx* c, uint32_t n) {
    uint32_t tiles = scan_tiles(n) + 1;
    if (c->scan_cap >= tiles) return PG_OK;
    c->scan_cap = 0;
    dfree(c->scan_tmp);
    PG_TRY(c, dalloc(&c->scan_tmp, tiles), "alloc scan");
    if (!c->d_total) PG_TRY(c, dalloc(&c->d_total, 1), "alloc scan total");
    c->scan_cap = tiles;
    return PG_OK;
}

// vertices of the per-path stacks a record pass may hold (128 B each): 2^22 paths of
// 13 vertices (max_depth 12), ~7 GB
static constexpr uint64_t kRecordVertexBudget = (1ull << 22) * 13;
// bounces (from 0) whose learned-guide queries are evaluated in Morton order, split by XCD
static constexpr uint32_t kInferSortDepth = 3;
static constexpr int kInferSortBits = 16;

static pg_status ensure_record_state(pg_ctx* c, uint32_t n_paths, uint32_t depth) {
    if (c->rs_paths >= n_paths && c->rs_depth >= depth) return PG_OK;
    // the sizes stay zero until every buffer is allocated, so a failed allocation
    // makes the next call allocate again instead of writing through null pointers
    c->rs_paths = 0;
    c->rs_depth = 0;
    dfree(c->rs.vpos); dfree(c->rs.vnrm); dfree(c->rs.vwi); dfree(c->rs.vwo); dfree(c->rs.vw); dfree(c->rs.ve);
    dfree(c->rs.vme); dfree(c->rs.vd); dfree(c->rs.nv);
    dfree(c->rec_counts); dfree(c->rec_offsets);
    size_t n = (size_t)n_paths * depth;
    PG_TRY(c, dalloc(&c->rs.vpos, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.vnrm, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.vwi, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.vwo, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.vw, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.ve, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.vme, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.vd, n), "alloc record vertices");
    PG_TRY(c, dalloc(&c->rs.nv, n_paths), "alloc record counts");
    PG_TRY(c, dalloc(&c->rec_counts, n_paths), "alloc record counts");
    PG_TRY(c, dalloc(&c->rec_offsets, n_paths), "alloc record offsets");
    c->rs.stride = n_paths;
    c->rs_paths = n_paths;
    c->rs_depth = depth;
    return ensure_scan(c, n_paths);
}

static pg_status ensure_record_capacity(pg_ctx* c, uint64_t need) {
    if (need <= c->rec_capacity) return PG_OK;
    uint64_t cap = c->rec_capacity ? c->rec_capacity : (1u << 20);
    while (cap < need) cap *= 2;
    if (cap > 0xffffffffull) return c->fail(PG_ERR_OVERFLOW, "record buffer exceeds 2^32 records");
    Record* nb = nullptr;
    TrainAux* na = nullptr;
    PG_TRY(c, dalloc(&nb, cap), "alloc records");
    pg_status st = c->hip(dalloc(&na, cap), "alloc record training data");
    if (st != PG_OK) {
        dfree(nb);
        return st;
    }
    if (c->records && c->n_records)
        PG_TRY(c, hipMemcpyAsync(nb, c->records, c->n_records * sizeof(Record), hipMemcpyDeviceToDevice, c->stream),
               "grow records");
    if (c->rec_aux && c->n_aux)
        PG_TRY(c, hipMemcpyAsync(na, c->rec_aux, c->n_aux * sizeof(TrainAux), hipMemcpyDeviceToDevice, c->stream),
               "grow records");
    PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
    dfree(c->records);
    dfree(c->rec_aux);
    c->records = nb;
    c->rec_aux = na;
    c->rec_capacity = cap;
    return PG_OK;
}

pg_status pg_render_spp(pg_ctx* c, const pg_render_params* params, uint32_t n_spp) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    pg_render_params prm = params ? *params : pg_render_params_default();
    const int integ = prm.integrator;
    if (integ != PG_INTEGRATOR_PATH && integ != PG_INTEGRATOR_RECORD && integ != PG_INTEGRATOR_GUIDED_QT &&
        integ != PG_INTEGRATOR_GUIDED_MLP && integ != PG_INTEGRATOR_GUIDED_PSS)
        return c->fail(PG_ERR_INVALID, "unknown integrator");
    if (prm.stack_lds % 16 != 0 || prm.stack_lds > 64)
        return c->fail(PG_ERR_INVALID, "stack_lds must be 0, 16, 32, 48 or 64");
    if ((integ == PG_INTEGRATOR_GUIDED_QT || integ == PG_INTEGRATOR_GUIDED_PSS) && !c->cache_ready)
        return c->fail(PG_ERR_STATE, "guided rendering needs pg_cache_build first");
    if (integ == PG_INTEGRATOR_GUIDED_MLP && !c->mlp.ready)
        return c->fail(PG_ERR_STATE, "learned guiding needs pg_mlp_init first");
    // shade mode: 0 BSDF sampling, 1 quadtree guide, 2 learned guide, 3 primary-sample-
    // space guide; a learned guide whose training diverged falls back to the quadtree
    // guide when a cache exists
    int mode = integ == PG_INTEGRATOR_GUIDED_QT ? 1 : (integ == PG_INTEGRATOR_GUIDED_PSS ? 3 : 0);
    if (integ == PG_INTEGRATOR_GUIDED_MLP) {
        if (!c->mlp.diverged) {
            mode = 2;
        } else {
            mode = c->cache_ready ? 1 : 0;
            c->mlp.fallbacks += n_spp;
        }
    }
    if (prm.max_depth > 64) return c->fail(PG_ERR_INVALID, "max_depth > 64");
    if (!(prm.guide_prob >= 0.f && prm.guide_prob <= 1.f)) return c->fail(PG_ERR_INVALID, "guide_prob outside [0,1]");
    if (prm.env_visibility > 1u) return c->fail(PG_ERR_INVALID, "env_visibility must be 0 or 1");
    if (integ == PG_INTEGRATOR_GUIDED_PSS && prm.knn_k != 0 && prm.knn_k != (uint32_t)kKnnMax)
        return c->fail(PG_ERR_INVALID, "primary-sample-space guide: knn_k must be 16 (exact) or 0 (approximate)");
    uint32_t W = (uint32_t)c->host.camera.width, H = (uint32_t)c->host.camera.height;
    uint32_t n = W * H;
    // record passes: the reference's record integrator, optionally guided by the learned guide
    const bool record_pass = integ == PG_INTEGRATOR_RECORD || (integ == PG_INTEGRATOR_GUIDED_MLP && prm.record);
    // samples traced together per pass: deep bounces hold few rays whose launches are
    // latency-bound, so several samples share them. Record passes keep per-path vertex
    // stacks of (max_depth + 1) x 128 B, so a record pass takes at most
    // kRecordVertexBudget / (max_depth + 1) paths (2^22 at depth 12, ~7 GB).
    uint32_t spp_pass = 1;
    if (!prm.sort_rays) {
        spp_pass = prm.samples_per_pass;
        const uint64_t record_paths =
            std::max<uint64_t>(1, std::min<uint64_t>(1ull << 22, kRecordVertexBudget / (prm.max_depth + 1)));
        const uint64_t max_paths = record_pass ? record_paths : (1ull << 26);
        if (record_pass && spp_pass > 1 && (uint64_t)spp_pass * n > max_paths)
            spp_pass = (uint32_t)std::max<uint64_t>(1, max_paths / n);
        if (spp_pass == 0) {
            // the call's samples split evenly over passes of at most max_paths paths
            const uint32_t cap = (uint32_t)std::max<uint64_t>(1, max_paths / n);
            const uint64_t passes = std::max<uint64_t>(1, ((uint64_t)n_spp + cap - 1) / cap);
            spp_pass = (uint32_t)std::max<uint64_t>(1, ((uint64_t)n_spp + passes - 1) / passes);
        }
        if ((uint64_t)spp_pass * n > 0x7fffffffull) return c->fail(PG_ERR_INVALID, "samples_per_pass x pixels >= 2^31");
    }
    const uint32_t np = n * spp_pass;  // paths of one pass
    pg_status st = ensure_paths(c, np);
    if (st != PG_OK) return st;
    if (record_pass && (st = ensure_record_state(c, np, prm.max_depth + 1)) != PG_OK) return st;
    c->rs.li_nee = prm.record_nee ? 1 : 0;
    if ((mode == 1 || mode == 3) && c->guide_cells_cap < np) {
        c->guide_cells_cap = 0;
        dfree(c->guide_cells);
        PG_TRY(c, dalloc(&c->guide_cells, (size_t)np * kKnnMax), "alloc guide cells");
        c->guide_cells_cap = np;
    }
    // sort buffers: kNN query order (quadtree guide) and ray trace order (sort_rays)
    if ((mode == 1 || mode == 2 || (mode == 3 && prm.knn_k != 0) || prm.sort_rays) && c->q_cap < np) {
        c->q_cap = 0;
        dfree(c->q_keys); dfree(c->q_vals); dfree(c->q_keys_tmp); dfree(c->q_vals_tmp); dfree(c->q_hist); dfree(c->q_scan);
        PG_TRY(c, dalloc(&c->q_keys, np), "alloc sort buffers");
        PG_TRY(c, dalloc(&c->q_vals, np), "alloc sort buffers");
        PG_TRY(c, dalloc(&c->q_keys_tmp, np), "alloc sort buffers");
        PG_TRY(c, dalloc(&c->q_vals_tmp, np), "alloc sort buffers");
        PG_TRY(c, dalloc(&c->q_hist, (size_t)256 * sort_tiles(np)), "alloc sort buffers");
        PG_TRY(c, dalloc(&c->q_scan, scan_tiles(256 * sort_tiles(np)) + 1), "alloc sort buffers");
        c->q_cap = np;
    }
    const float3 q_lo = make_float3(c->host.centre[0] - c->host.radius, c->host.centre[1] - c->host.radius,
                                    c->host.centre[2] - c->host.radius);
    const float q_inv = c->host.radius > 0.f ? 1.f / (2.f * c->host.radius) : 0.f;
    // learned guide: the bounce's queries in Morton order of their shading points, for
    // the first bounces (whose queues fill most of the pass; deeper queues are short and
    // would pay for sorting the whole capacity)
    auto infer_order = [&](int parity, uint32_t depth, const uint32_t*& order) -> hipError_t {
        order = nullptr;
        if (depth > kInferSortDepth) return hipSuccess;
        hipError_t e = launch_query_keys(c->dev, c->ps, q_lo, q_inv, parity, np, c->q_keys, c->q_vals, c->stream,
                                         kInferSortBits);
        if (e == hipSuccess)
            e = radix_sort_pairs(c->q_keys, c->q_vals, c->q_keys_tmp, c->q_vals_tmp, np, kInferSortBits, c->q_hist,
                                 c->q_scan, c->stream);
        order = c->q_vals;
        return e;
    };
    // trace order of a ray queue (shadow = 1: the shadow queue) for sort_rays
    auto ray_order = [&](int parity, int shadow) -> hipError_t {
        hipError_t e = launch_ray_keys(c->ps, q_lo, q_inv, parity, shadow, np, c->q_keys, c->q_vals, c->stream);
        if (e != hipSuccess) return e;
        return radix_sort_pairs(c->q_keys, c->q_vals, c->q_keys_tmp, c->q_vals_tmp, np, 24, c->q_hist, c->q_scan,
                                c->stream);
    };
    const bool sort_rays = prm.sort_rays != 0 && !record_pass;
    const bool dyn = prm.dyn_fetch != 0 && !sort_rays;
    if (mode == 2 && c->mix_cap < np) {
        c->mix_cap = 0;
        dfree(c->mix);
        PG_TRY(c, dalloc(&c->mix, (size_t)np * kMixF4), "alloc guide mixtures");
        c->mix_cap = np;
    }
    const int stack = lds_entries((int)c->host.bvh.max_depth, prm.stack_lds, c->ps);
    const uint32_t grid_trace = prm.trace_blocks_per_cu ? (uint32_t)c->cu_count * prm.trace_blocks_per_cu
                                                        : trace_grid(c->cu_count, stack, dyn);
    const uint32_t grid_shade = (uint32_t)c->cu_count * (prm.shade_blocks_per_cu ? prm.shade_blocks_per_cu : kShadeBlocksPerCu);
    const uint32_t grid_guide = (uint32_t)c->cu_count * (c->gv.levels <= 16 ? 10u : 5u);
    const uint32_t grid_mlp = (uint32_t)c->cu_count * 3;  // 48 KB of LDS per block: 3 blocks per CU
    GuideCtx gc{c->gv, (mode == 1 || mode == 3) ? c->guide_cells : nullptr, mode == 2 ? c->mix : nullptr};
    gc.g.guide_prob = prm.guide_prob;
    c->pool_used = 0;
    c->launches.clear();
    auto timed = [&](int kind, auto&& fn) -> hipError_t {
        hipEvent_t a = nullptr, b = nullptr;
        if (c->timing) {
            a = c->ev();
            b = c->ev();
            if (a && b) (void)hipEventRecord(a, c->stream);
        }
        hipError_t e = fn();
        if (c->timing && a && b) {
            (void)hipEventRecord(b, c->stream);
            c->launches.push_back(TimedLaunch{kind, a, b});
        }
        return e;
    };
    hipEvent_t t0 = c->ev(), t1 = c->ev();
    (void)hipEventRecord(t0, c->stream);
    for (uint32_t s = 0; s < n_spp;) {
        const uint32_t ns = std::min(spp_pass, n_spp - s);
        s += ns;
        PassParams pp{prm.seed,           c->spp, W, H, prm.max_depth, prm.rr_depth, prm.integrator, prm.env_visibility,
                      ns,                 prm.guide_prob};
        PG_TRY(c, timed(3, [&] { return launch_raygen(c->dev, c->ps, pp, c->stream); }), "raygen");
        for (uint32_t depth = 0; depth <= prm.max_depth; ++depth) {
            int parity = (int)(depth & 1u);
            // guided bounces; deeper ones sample the BSDF alone (the MIS at every vertex
            // uses that vertex's own density, so mixing the two keeps the estimate unbiased)
            const bool guide_here = prm.guide_depth == 0 || depth < prm.guide_depth;
            const int mode_here = guide_here ? mode : 0;
            const bool sort_here = sort_rays && depth > 0;
            if (sort_here) PG_TRY(c, timed(3, [&] { return ray_order(parity, 0); }), "ray order");
            PG_TRY(c, timed(0, [&] { return launch_extend(c->dev, c->ps, parity, stack, c->counting, grid_trace, c->stream, false, sort_here ? c->q_vals : nullptr, dyn); }), "extend");
            c->n_extend++;
            if (record_pass) {
                if (mode_here == 2) {
                    const uint32_t* ord = nullptr;
                    PG_TRY(c, timed(4, [&] { return infer_order(parity, depth, ord); }), "query order");
                    PG_TRY(c, timed(4, [&] { return launch_mlp_infer(c->dev, c->ps, c->mlp.view, parity, c->mix, grid_mlp, c->stream, false, 0u, ord); }), "mlp_infer");
                }
                const float4* mx = mode_here == 2 ? c->mix : nullptr;
                PG_TRY(c, timed(1, [&] { return launch_shade_record(c->dev, c->ps, pp, c->rs, mx, parity, (int)depth, grid_shade, c->stream); }), "shade_record");
                if (c->rs.li_nee) {
                    PG_TRY(c, timed(2, [&] { return launch_shadow_record(c->dev, c->ps, c->rs, stack, grid_trace, c->stream); }), "shadow_record");
                    c->n_shadow++;
                }
                continue;
            }
            if (mode_here == 3 && prm.knn_k == 0) {
                // approximate neighbours: no query sort needed (no search stack)
                PG_TRY(c, timed(4, [&] { return launch_guide_cells(c->dev, c->ps, c->gv, c->guide_cells, parity, grid_guide, c->stream, nullptr, 0, 2); }), "guide records");
            } else if (mode_here == 3) {
                PG_TRY(c, timed(4, [&] {
                    hipError_t e = launch_query_keys(c->dev, c->ps, q_lo, q_inv, parity, np, c->q_keys, c->q_vals, c->stream);
                    if (e == hipSuccess)
                        e = radix_sort_pairs(c->q_keys, c->q_vals, c->q_keys_tmp, c->q_vals_tmp, np, 24, c->q_hist,
                                             c->q_scan, c->stream);
                    return e;
                }), "guide query order");
                PG_TRY(c, timed(4, [&] { return launch_guide_cells(c->dev, c->ps, c->gv, c->guide_cells, parity, grid_guide, c->stream, c->q_vals, np, 1); }), "guide records");
            }
            if (mode_here == 1) {
                PG_TRY(c, timed(4, [&] {
                    hipError_t e = launch_query_keys(c->dev, c->ps, q_lo, q_inv, parity, np, c->q_keys, c->q_vals, c->stream);
                    if (e == hipSuccess)
                        e = radix_sort_pairs(c->q_keys, c->q_vals, c->q_keys_tmp, c->q_vals_tmp, np, 24, c->q_hist,
                                             c->q_scan, c->stream);
                    return e;
                }), "guide query order");
                PG_TRY(c, timed(4, [&] { return launch_guide_cells(c->dev, c->ps, c->gv, c->guide_cells, parity, grid_guide, c->stream, c->q_vals, np, 0); }), "guide_cells");
            }
            if (mode_here == 2) {
                const uint32_t* ord = nullptr;
                PG_TRY(c, timed(4, [&] { return infer_order(parity, depth, ord); }), "query order");
                PG_TRY(c, timed(4, [&] { return launch_mlp_infer(c->dev, c->ps, c->mlp.view, parity, c->mix, grid_mlp, c->stream, false, 0u, ord); }), "mlp_infer");
            }
            PG_TRY(c, timed(1, [&] { return launch_shade(c->dev, c->ps, pp, gc, mode_here, parity, (int)depth, grid_shade, c->stream); }), "shade");
            if (sort_rays) PG_TRY(c, timed(3, [&] { return ray_order(parity, 1); }), "shadow order");
            PG_TRY(c, timed(2, [&] { return launch_shadow(c->dev, c->ps, stack, c->counting, grid_trace, c->stream, false, sort_rays ? c->q_vals : nullptr, dyn); }), "shadow");
            c->n_shadow++;
        }
        if (record_pass) {
            // back-propagate, then append this pass's records in (path, depth) order: path
            // i is sample i / n of pixel i % n, the order of one-sample passes
            const uint32_t npp = n * ns;
            PG_TRY(c, launch_backprop_count(c->rs, npp, c->rec_counts, c->stream), "backprop count");
            PG_TRY(c, exclusive_scan(c->rec_counts, c->rec_offsets, npp, c->scan_tmp, c->d_total, c->stream), "scan");
            uint32_t total = 0;
            PG_TRY(c, hipMemcpyAsync(&total, c->d_total, 4, hipMemcpyDeviceToHost, c->stream), "read total");
            PG_TRY(c, hipStreamSynchronize(c->stream), "sync");
            uint32_t cap = 0;
            uint64_t base = c->n_records;
            if (prm.record) {
                st = ensure_record_capacity(c, base + total);
                if (st != PG_OK) return st;
                cap = (uint32_t)std::min<uint64_t>(c->rec_capacity, 0xffffffffull);
            }
            TrainAux* aux = (prm.record && c->n_aux == c->n_records) ? c->rec_aux : nullptr;
            PG_TRY(c, launch_backprop_write(c->rs, npp, c->rec_offsets, (uint32_t)base, c->records, aux, cap,
                                            c->ps.radiance, c->stream), "backprop write");
            if (prm.record) {
                if (aux) c->n_aux += total;
                c->n_records += total;
                c->rec_gen++;
            }
            c->cache_ready = false;
        }
        PG_TRY(c, timed(3, [&] { return launch_resolve(c->ps, pp, c->film, c->stream); }), "resolve");
        c->spp += ns;
        c->paths += (uint64_t)n * ns;
    }
    (void)hipEventRecord(t1, c->stream);
    PG_TRY(c, hipStreamSynchronize(c->stream), "render");
    float tot = 0.f;
    (void)hipEventElapsedTime(&tot, t0, t1);
    c->ms_total += tot;
    for (const TimedLaunch& l : c->launches) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) c->ms[l.kind] += ms;
    }
    return PG_OK;
}

pg_status pg_render_tune(pg_ctx* c, const pg_render_params* params, pg_render_params* tuned, double* ms_out) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    pg_render_params best = params ? *params : pg_render_params_default();
    if (best.integrator == PG_INTEGRATOR_RECORD || (best.integrator == PG_INTEGRATOR_GUIDED_MLP && best.record))
        return c->fail(PG_ERR_INVALID, "record passes are not tuned");
    if (best.trace_blocks_per_cu > 32 || best.shade_blocks_per_cu > 32) return c->fail(PG_ERR_INVALID, "blocks per CU");
    // each candidate renders whole passes of the caller's samples_per_pass (0: 1), the
    // shape its renders will have: a warm-up pass, then two timed passes (one from 8 SPP)
    const uint32_t k = best.samples_per_pass ? best.samples_per_pass : 1u;
    const uint32_t timed_spp = k >= 8 ? k : 2 * k;
    auto measure = [&](const pg_render_params& p, double& ms) -> pg_status {
        pg_status st = pg_film_clear(c);
        if (st != PG_OK) return st;
        if ((st = pg_render_spp(c, &p, k)) != PG_OK) return st;  // warm-up
        double before = c->ms_total;
        if ((st = pg_render_spp(c, &p, timed_spp)) != PG_OK) return st;
        ms = (c->ms_total - before) / timed_spp;
        return PG_OK;
    };
    double best_ms = 0.0;
    pg_status st = measure(best, best_ms);
    if (st != PG_OK) return st;
    const uint32_t lds_cand[] = {0u, 32u, 64u};
    const uint32_t trace_cand[] = {0u, 8u, 10u, 12u, 14u, 16u};
    const uint32_t shade_cand[] = {2u, 3u, 4u, 6u};
    {
        pg_render_params p = best;
        p.dyn_fetch = best.dyn_fetch ? 0u : 1u;
        double ms = 0.0;
        if ((st = measure(p, ms)) != PG_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = p; }
    }
    for (uint32_t v : lds_cand) {
        pg_render_params p = best;
        if (p.stack_lds == v) continue;
        p.stack_lds = v;
        double ms = 0.0;
        if ((st = measure(p, ms)) != PG_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = p; }
    }
    for (uint32_t v : trace_cand) {
        pg_render_params p = best;
        if (p.trace_blocks_per_cu == v) continue;
        p.trace_blocks_per_cu = v;
        double ms = 0.0;
        if ((st = measure(p, ms)) != PG_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = p; }
    }
    for (uint32_t v : shade_cand) {
        pg_render_params p = best;
        if ((p.shade_blocks_per_cu ? p.shade_blocks_per_cu : kShadeBlocksPerCu) == v) continue;
        p.shade_blocks_per_cu = v;
        double ms = 0.0;
        if ((st = measure(p, ms)) != PG_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = p; }
    }
    {
        pg_render_params p = best;
        p.sort_rays = best.sort_rays ? 0u : 1u;
        double ms = 0.0;
        if ((st = measure(p, ms)) != PG_OK) return st;
        if (ms < best_ms) { best_ms = ms; best = p; }
    }
    if ((st = pg_film_clear(c)) != PG_OK) return st;
    if (tuned) *tuned = best;
    if (ms_out) *ms_out = best_ms;
    return PG_OK;
}

// ------------------------------------------------------------------ records and cache
pg_status pg_records_count(pg_ctx* c, uint64_t* n) {
    PG_CHECK_CTX(c);
    if (n) *n = c->n_records;
    return PG_OK;
}

pg_status pg_records_read(pg_ctx* c, pg_path_vertex* out, uint64_t cap, uint64_t* n) {
    PG_CHECK_CTX(c);
    uint64_t m = std::min<uint64_t>(cap, c->n_records);
    if (m && !out) return c->fail(PG_ERR_INVALID, "null output");
    if (m) PG_TRY(c, hipMemcpy(out, c->records, m * sizeof(Record), hipMemcpyDeviceToHost), "read records");
    if (n) *n = m;
    return PG_OK;
}

pg_status pg_records_clear(pg_ctx* c) {
    PG_CHECK_CTX(c);
    c->n_records = 0;
    c->n_aux = 0;
    c->rec_gen++;
    c->cache_ready = false;
    return PG_OK;
}

pg_status pg_records_upload(pg_ctx* c, const pg_path_vertex* in, uint64_t n) {
    PG_CHECK_CTX(c);
    if (n && !in) return c->fail(PG_ERR_INVALID, "null records");
    pg_status st = ensure_record_capacity(c, n);
    if (st != PG_OK) return st;
    if (n) PG_TRY(c, hipMemcpy(c->records, in, n * sizeof(Record), hipMemcpyHostToDevice), "upload records");
    c->n_records = n;
    c->n_aux = 0;
    c->rec_gen++;
    c->cache_ready = false;
    return PG_OK;
}

pg_status pg_cache_build(pg_ctx* c, uint32_t cell_size, uint32_t qt_nodes_per_cell) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (cell_size == 0) cell_size = 4096;
    if (qt_nodes_per_cell == 0) qt_nodes_per_cell = 1024;
    if (cell_size < kLeafRecords || qt_nodes_per_cell < 1)
        return c->fail(PG_ERR_INVALID, "cell size / quadtree budget");
    if (c->n_records > 0x7fffffffull) return c->fail(PG_ERR_OVERFLOW, "too many records");
    uint32_t n = (uint32_t)c->n_records;
    uint32_t leaves = (n + kLeafRecords - 1) / kLeafRecords;
    uint32_t P = 1, levels = 0;
    while (P < leaves) { P <<= 1; ++levels; }
    uint32_t cells = (n + cell_size - 1) / cell_size;
    if (cells >= (1u << 24)) return c->fail(PG_ERR_OVERFLOW, "more than 2^24 guide cells: raise cell_size");
    hipEvent_t e0 = c->ev(), e1 = c->ev();
    if (n > c->cache_cap || P > c->cache_pow2_cap || cells * qt_nodes_per_cell > c->cache_cells_cap * c->cache_stride ||
        !c->cb.keys) {
        c->free_cache();
        uint32_t cap = std::max<uint32_t>(n, 1024u);
        uint32_t tiles = sort_tiles(cap);
        PG_TRY(c, dalloc(&c->cb.keys, cap), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.vals, cap), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.keys_tmp, cap), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.vals_tmp, cap), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.hist, (size_t)256 * tiles), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.scan_tmp, scan_tiles(256 * tiles) + 1), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.pos, cap), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.dir, cap), "alloc cache");
        PG_TRY(c, dalloc(&c->cb.nodes, (size_t)2 * (2 * P - 1)), "alloc cache");
        uint32_t cells_cap = std::max<uint32_t>(cells, 1u);
        PG_TRY(c, dalloc(&c->cb.qt, (size_t)cells_cap * qt_nodes_per_cell), "alloc quadtrees");
        PG_TRY(c, dalloc(&c->cb.qt_acc, (size_t)cells_cap * qt_nodes_per_cell * 4), "alloc quadtrees");
        c->cache_cap = cap;
        c->cache_pow2_cap = P;
        c->cache_cells_cap = cells_cap;
        c->cache_stride = qt_nodes_per_cell;
    }
    c->cb.leaf_pow2 = P;
    c->cb.levels = levels;
    // Morton frame: the scene bounds contain every record position
    const PreparedScene& h = c->host;
    float ext = 2.f * h.radius;
    float3 lo = make_float3(h.centre[0] - h.radius, h.centre[1] - h.radius, h.centre[2] - h.radius);
    float inv = ext > 0.f ? 1.f / ext : 0.f;
    (void)hipEventRecord(e0, c->stream);
    if (n) PG_TRY(c, launch_cache_build(c->records, n, lo, make_float3(inv, inv, inv), c->cb, cell_size,
                                        qt_nodes_per_cell, c->stream), "cache build");
    (void)hipEventRecord(e1, c->stream);
    PG_TRY(c, hipStreamSynchronize(c->stream), "cache build");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c->cache_build_ms = ms;
    c->pool_used = 0;
    GuideView& g = c->gv;
    g.n_records = n;
    g.n_leaf_pow2 = P;
    g.levels = levels;
    g.cell_size = cell_size;
    g.n_cells = cells;
    g.qt_stride = qt_nodes_per_cell;
    g.rec_pos = c->cb.pos;
    g.nodes = c->cb.nodes;
    g.qt = c->cb.qt;
    g.rec_dir = c->cb.dir;
    g.rec_key = c->cb.keys;
    g.m_lo[0] = lo.x; g.m_lo[1] = lo.y; g.m_lo[2] = lo.z;
    g.m_inv = inv;
    g.guide_prob = 0.5f;
    // neighbour lists of the leaves (approximate neighbours of the PSS guide)
    const uint32_t n_leaves = (n + kLeafRecords - 1) / kLeafRecords;
    if (n_leaves > c->leaf_nbrs_cap) {
        c->leaf_nbrs_cap = 0;
        dfree(c->leaf_nbrs);
        PG_TRY(c, dalloc(&c->leaf_nbrs, (size_t)n_leaves * kKnnMax), "alloc leaf neighbours");
        c->leaf_nbrs_cap = n_leaves;
    }
    g.leaf_nbrs = c->leaf_nbrs;
    if (n) {
        PG_TRY(c, launch_leaf_nbrs(g, n_leaves, c->leaf_nbrs, c->stream), "leaf neighbours");
        PG_TRY(c, hipStreamSynchronize(c->stream), "leaf neighbours");
    }
    c->cache_ready = true;
    return PG_OK;
}

pg_status pg_cache_info(pg_ctx* c, uint64_t* n_records, uint32_t* n_nodes, uint32_t* depth, uint32_t* n_cells,
                        double* build_ms) {
    PG_CHECK_CTX(c);
    if (!c->cache_ready) return c->fail(PG_ERR_STATE, "cache not built");
    if (n_records) *n_records = c->gv.n_records;
    if (n_nodes) *n_nodes = 2 * c->gv.n_leaf_pow2 - 1;
    if (depth) *depth = c->gv.levels;
    if (n_cells) *n_cells = c->gv.n_cells;
    if (build_ms) *build_ms = c->cache_build_ms;
    return PG_OK;
}

pg_status pg_cache_export(pg_ctx* c, float* nodes, float* sorted_pos, float* quadtrees) {
    PG_CHECK_CTX(c);
    if (!c->cache_ready) return c->fail(PG_ERR_STATE, "cache not built");
    const GuideView& g = c->gv;
    if (nodes) PG_TRY(c, hipMemcpy(nodes, g.nodes, (size_t)(2 * g.n_leaf_pow2 - 1) * 32, hipMemcpyDeviceToHost), "export");
    if (sorted_pos && g.n_records)
        PG_TRY(c, hipMemcpy(sorted_pos, g.rec_pos, (size_t)g.n_records * 16, hipMemcpyDeviceToHost), "export");
    if (quadtrees && g.n_cells)
        PG_TRY(c, hipMemcpy(quadtrees, g.qt, (size_t)g.n_cells * g.qt_stride * sizeof(QtNode), hipMemcpyDeviceToHost),
               "export");
    return PG_OK;
}

pg_status pg_knn(pg_ctx* c, const float* q, uint32_t nq, uint32_t k, uint32_t* idx, float* d2) {
    PG_CHECK_CTX(c);
    if (!c->cache_ready) return c->fail(PG_ERR_STATE, "cache not built");
    if (k == 0 || k > (uint32_t)kKnnMax) return c->fail(PG_ERR_INVALID, "k must be in 1..16");
    if (nq == 0) return PG_OK;
    float *dq = nullptr, *dd = nullptr;
    uint32_t* di = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dq, (size_t)nq * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&di, (size_t)nq * k), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dd, (size_t)nq * k), "alloc")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(dq, q, (size_t)nq * 12, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(launch_knn(c->gv, dq, nq, k, di, dd, c->stream), "knn")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        hipMemcpy(idx, di, (size_t)nq * k * 4, hipMemcpyDeviceToHost);
        st = c->hip(hipMemcpy(d2, dd, (size_t)nq * k * 4, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(dq); dfree(di); dfree(dd);
    return st;
}

pg_status pg_guide_qt_eval(pg_ctx* c, const float* pos, const float* u3, const float* wi_eval, uint32_t n,
                           float* wi_out, float* pdf_sample, float* pdf_eval) {
    PG_CHECK_CTX(c);
    if (!c->cache_ready) return c->fail(PG_ERR_STATE, "cache not built");
    if (n == 0) return PG_OK;
    float *dp = nullptr, *du = nullptr, *dw = nullptr, *dwo = nullptr, *dps = nullptr, *dpe = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dp, (size_t)n * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&du, (size_t)n * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dw, (size_t)n * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dwo, (size_t)n * 3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dps, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dpe, n), "alloc")) != PG_OK) break;
        hipMemcpy(dp, pos, (size_t)n * 12, hipMemcpyHostToDevice);
        hipMemcpy(du, u3, (size_t)n * 12, hipMemcpyHostToDevice);
        if ((st = c->hip(hipMemcpy(dw, wi_eval, (size_t)n * 12, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(launch_guide_eval(c->gv, dp, du, dw, n, dwo, dps, dpe, c->stream), "guide eval")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        hipMemcpy(wi_out, dwo, (size_t)n * 12, hipMemcpyDeviceToHost);
        hipMemcpy(pdf_sample, dps, (size_t)n * 4, hipMemcpyDeviceToHost);
        st = c->hip(hipMemcpy(pdf_eval, dpe, (size_t)n * 4, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(dp); dfree(du); dfree(dw); dfree(dwo); dfree(dps); dfree(dpe);
    return st;
}

pg_status pg_guide_pss_eval(pg_ctx* c, uint32_t material, const float* pos, const float* nrm, const float* wo,
                            const float* u3, const float* wi_eval, uint32_t n, float* wi_out, float* pdf_sample,
                            float* pdf_eval) {
    PG_CHECK_CTX(c);
    if (!c->cache_ready) return c->fail(PG_ERR_STATE, "cache not built");
    if (material >= c->host.mats.size()) return c->fail(PG_ERR_INVALID, "unknown material");
    if (n == 0) return PG_OK;
    if (!pos || !nrm || !wo || !u3 || !wi_eval || !wi_out || !pdf_sample || !pdf_eval)
        return c->fail(PG_ERR_INVALID, "null array");
    float *din = nullptr, *dout = nullptr;
    pg_status st = PG_OK;
    const size_t v3 = (size_t)n * 3;
    do {
        // inputs: pos, nrm, wo, u3, wi_eval; outputs: wi_out, pdf_sample, pdf_eval
        if ((st = c->hip(dalloc(&din, 5 * v3), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dout, v3 + 2 * (size_t)n), "alloc")) != PG_OK) break;
        const float* src[5] = {pos, nrm, wo, u3, wi_eval};
        for (int k = 0; k < 5 && st == PG_OK; ++k)
            st = c->hip(hipMemcpy(din + k * v3, src[k], v3 * 4, hipMemcpyHostToDevice), "copy");
        if (st != PG_OK) break;
        if ((st = c->hip(launch_pss_eval(c->dev, c->gv, material, din, din + v3, din + 2 * v3, din + 3 * v3,
                                         din + 4 * v3, n, dout, dout + v3, dout + v3 + n, c->stream),
                         "pss eval")) != PG_OK)
            break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(wi_out, dout, v3 * 4, hipMemcpyDeviceToHost), "copy")) != PG_OK) break;
        if ((st = c->hip(hipMemcpy(pdf_sample, dout + v3, (size_t)n * 4, hipMemcpyDeviceToHost), "copy")) != PG_OK) break;
        st = c->hip(hipMemcpy(pdf_eval, dout + v3 + n, (size_t)n * 4, hipMemcpyDeviceToHost), "copy");
    } while (0);
    dfree(din);
    dfree(dout);
    return st;
}

pg_status pg_film_read(pg_ctx* c, float* rgb_sum, uint32_t* spp) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (rgb_sum) {
        std::vector<float4> tmp(c->film_n);
        PG_TRY(c, hipMemcpy(tmp.data(), c->film, (size_t)c->film_n * 16, hipMemcpyDeviceToHost), "read film");
        for (uint32_t i = 0; i < c->film_n; ++i) {
            rgb_sum[3 * i] = tmp[i].x; rgb_sum[3 * i + 1] = tmp[i].y; rgb_sum[3 * i + 2] = tmp[i].z;
        }
    }
    if (spp) *spp = c->spp;
    return PG_OK;
}

// Film::tonemap (Imaging.h:220-244) of every pixel of the film.
__global__ void k_tonemap(const float4* __restrict__ film, uint32_t n, float inv_spp, float exposure, float lw2,
                          uint8_t* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 f = film[i];
    float cr = (f.x * inv_spp) * exposure, cg = (f.y * inv_spp) * exposure, cb = (f.z * inv_spp) * exposure;
    float lin = ((0.2126f * cr) + (0.7152f * cg)) + (0.0722f * cb);
    uint8_t r = 0, g = 0, b = 0;
    if (!(lin < 1e-4f)) {
        float lout = lin * (1.f + (lin / lw2)) / (1.f + lin);
        float k = lout / lin;
        const float ge = 1.f / 2.2f;
        r = (uint8_t)(pg::smax(0.f, pg::smin(powf(cr * k, ge), 1.f)) * 255.f);
        g = (uint8_t)(pg::smax(0.f, pg::smin(powf(cg * k, ge), 1.f)) * 255.f);
        b = (uint8_t)(pg::smax(0.f, pg::smin(powf(cb * k, ge), 1.f)) * 255.f);
    }
    out[3 * (size_t)i] = r;
    out[3 * (size_t)i + 1] = g;
    out[3 * (size_t)i + 2] = b;
}

pg_status pg_film_tonemap(pg_ctx* c, float exposure, float lmax_white, uint8_t* rgb8) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    if (!rgb8) return c->fail(PG_ERR_INVALID, "null output");
    uint32_t n = c->film_n;
    uint8_t* d = nullptr;
    PG_TRY(c, dalloc(&d, (size_t)n * 3), "alloc tonemap output");
    float inv = c->spp > 0 ? 1.f / (float)c->spp : 1.f;
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->film, n, inv, exposure,
                       lmax_white * lmax_white, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(rgb8, d, (size_t)n * 3, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d);
    if (e != hipSuccess) return c->fail(PG_ERR_DEVICE, std::string("tonemap: ") + hipGetErrorString(e));
    return PG_OK;
}

pg_status pg_stats_read(pg_ctx* c, pg_stats* s) {
    PG_CHECK_CTX(c);
    if (!s) return c->fail(PG_ERR_INVALID, "null stats");
    unsigned long long h[8];
    PG_TRY(c, hipMemcpy(h, c->d_stats, sizeof(h), hipMemcpyDeviceToHost), "read stats");
    std::memset(s, 0, sizeof(*s));
    s->closest_rays = h[0];
    s->shadow_rays = h[1];
    s->nodes_visited = h[2];
    s->tris_tested = h[3];
    s->shadow_nodes_visited = h[4];
    s->shadow_tris_tested = h[5];
    s->paths = c->paths;
    s->kernel_ms_extend = c->ms[0];
    s->kernel_ms_shade = c->ms[1];
    s->kernel_ms_shadow = c->ms[2];
    s->kernel_ms_guide = c->ms[4];
    s->records = c->n_records;
    s->kernel_ms_total = c->ms_total;
    s->extend_launches = c->n_extend;
    s->shadow_launches = c->n_shadow;
    return PG_OK;
}

pg_status pg_stats_reset(pg_ctx* c) {
    PG_CHECK_CTX(c);
    PG_TRY(c, hipMemset(c->d_stats, 0, 8 * sizeof(unsigned long long)), "reset stats");
    for (double& m : c->ms) m = 0;
    c->ms_total = 0;
    c->n_extend = c->n_shadow = c->paths = 0;
    return PG_OK;
}

pg_status pg_copy_benchmark(pg_ctx* c, uint64_t bytes, uint32_t iters, double* ms_per_copy) {
    PG_CHECK_CTX(c);
    if (bytes < 16 || iters == 0) return c->fail(PG_ERR_INVALID, "copy size / iterations");
    size_t n = (size_t)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    pg_status st = PG_OK;
    hipEvent_t e0 = c->ev(), e1 = c->ev();
    do {
        if ((st = c->hip(dalloc(&a, n), "alloc copy src")) != PG_OK) break;
        if ((st = c->hip(dalloc(&b, n), "alloc copy dst")) != PG_OK) break;
        if ((st = c->hip(hipMemsetAsync(a, 0, n * 16, c->stream), "init copy src")) != PG_OK) break;
        if ((st = c->hip(launch_copy(a, b, n, c->cu_count, c->stream), "copy warmup")) != PG_OK) break;
        (void)hipEventRecord(e0, c->stream);
        for (uint32_t i = 0; i < iters && st == PG_OK; ++i)
            st = c->hip(launch_copy(a, b, n, c->cu_count, c->stream), "copy");
        (void)hipEventRecord(e1, c->stream);
        if (st != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms_per_copy) *ms_per_copy = ms / iters;
    } while (0);
    c->pool_used = 0;
    dfree(a);
    dfree(b);
    return st;
}

pg_status pg_set_timing(pg_ctx* c, int32_t enable) {
    PG_CHECK_CTX(c);
    c->timing = enable != 0;
    return PG_OK;
}

pg_status pg_set_counting(pg_ctx* c, int32_t enable) {
    PG_CHECK_CTX(c);
    c->counting = enable != 0;
    return PG_OK;
}


// ------------------------------------------------------------------ learned guide
pg_mlp_config pg_mlp_config_default(void) {
    pg_mlp_config m;
    std::memset(&m, 0, sizeof(m));
    m.seed = 1;
    m.log2_table = 19;
    m.res_min = 16.f;
    m.res_max = 1024.f;
    m.lr = 1e-2f;
    m.beta1 = 0.9f;
    m.beta2 = 0.99f;
    m.eps = 1e-15f;
    m.batch = 1u << 18;
    m.weight_clip = 100.f;
    m.act_budget = 512ull << 20;
    m.nan_step = 0;
    return m;
}

static uint32_t mlp_chunk(const pg_mlp_config& cfg, uint32_t n) {
    uint64_t rows = cfg.act_budget / (uint64_t)(kRowLen * 2);
    rows &= ~127ull;
    if (rows < 128) rows = 128;
    return (uint32_t)std::min<uint64_t>(rows, n);
}

static pg_status mlp_buffers(pg_ctx* c, uint32_t batch, uint32_t steps) {
    auto& m = c->mlp;
    const uint32_t chunk = mlp_chunk(m.cfg, batch);
    const uint32_t groups = mlp_wgrad_groups(chunk);
    if (m.batch_cap < batch) {
        m.batch_cap = m.loss_cap = 0;
        dfree(m.batch); dfree(m.loss); dfree(m.batch_raw);
        dfree(m.sort.keys); dfree(m.sort.vals); dfree(m.sort.keys_tmp); dfree(m.sort.vals_tmp);
        dfree(m.sort.hist); dfree(m.sort.scan);
        PG_TRY(c, dalloc(&m.batch, batch), "alloc training batch");
        PG_TRY(c, dalloc(&m.loss, batch), "alloc training batch");
        PG_TRY(c, dalloc(&m.batch_raw, batch), "alloc training batch");
        PG_TRY(c, dalloc(&m.sort.keys, batch), "alloc batch sort");
        PG_TRY(c, dalloc(&m.sort.vals, batch), "alloc batch sort");
        PG_TRY(c, dalloc(&m.sort.keys_tmp, batch), "alloc batch sort");
        PG_TRY(c, dalloc(&m.sort.vals_tmp, batch), "alloc batch sort");
        PG_TRY(c, dalloc(&m.sort.hist, (size_t)256 * sort_tiles(batch)), "alloc batch sort");
        PG_TRY(c, dalloc(&m.sort.scan, scan_tiles(256 * sort_tiles(batch)) + 1), "alloc batch sort");
        m.batch_cap = m.loss_cap = batch;
    }
    if (m.rows_cap < chunk) {
        m.rows_cap = 0;
        dfree(m.rows);
        PG_TRY(c, dalloc(&m.rows, (size_t)chunk * kRowLen), "alloc activation rows");
        m.rows_cap = chunk;
    }
    if (m.groups_cap < groups) {
        m.groups_cap = 0;
        dfree(m.partial);
        PG_TRY(c, dalloc(&m.partial, (size_t)groups * kParams), "alloc gradient partials");
        m.groups_cap = groups;
    }
    if (m.scal_cap < steps + 1) {
        m.scal_cap = 0;
        dfree(m.scal);
        PG_TRY(c, dalloc(&m.scal, steps + 1), "alloc loss history");
        m.scal_cap = steps + 1;
    }
    return PG_OK;
}

// Forward/backward of `batch` samples already in m.batch: per chunk the fused
// kernel then the weight-gradient GEMM (partials accumulate over chunks).
static pg_status mlp_backward(pg_ctx* c, uint32_t batch, int nan_inject, uint32_t* groups_out) {
    auto& m = c->mlp;
    const uint32_t chunk = mlp_chunk(m.cfg, batch);
    const uint32_t groups = mlp_wgrad_groups(chunk);
    uint32_t k = 0;
    for (uint32_t s0 = 0; s0 < batch; s0 += chunk, ++k) {
        uint32_t cn = std::min(chunk, batch - s0);
        PG_TRY(c, launch_mlp_train_chunk(m.view, m.batch + s0, cn, 1.f / (float)batch, m.rows, m.loss + s0, m.ggrad,
                                         (nan_inject && k == 0) ? 1 : 0, c->cu_count, c->stream), "mlp train");
        PG_TRY(c, launch_mlp_wgrad(m.rows, cn, groups, m.partial, k > 0 ? 1 : 0, c->stream), "mlp wgrad");
    }
    m.chunks = k;
    *groups_out = groups;
    return PG_OK;
}

pg_status pg_mlp_init(pg_ctx* c, const pg_mlp_config* cfg_in) {
    PG_CHECK_CTX(c);
    if (!c->has_scene) return c->fail(PG_ERR_STATE, "no scene uploaded");
    pg_mlp_config cfg = cfg_in ? *cfg_in : pg_mlp_config_default();
    if (cfg.log2_table < 12 || cfg.log2_table > 22) return c->fail(PG_ERR_INVALID, "log2_table outside 12..22");
    if (!(cfg.res_min >= 1.f) || !(cfg.res_max >= cfg.res_min) || !(cfg.res_max <= 1048576.f))
        return c->fail(PG_ERR_INVALID, "grid resolutions");
    if (!(cfg.lr > 0.f) || !(cfg.beta1 >= 0.f && cfg.beta1 < 1.f) || !(cfg.beta2 >= 0.f && cfg.beta2 < 1.f) ||
        !(cfg.eps >= 0.f))
        return c->fail(PG_ERR_INVALID, "Adam parameters");
    if (cfg.batch == 0 || cfg.batch > (1u << 24)) return c->fail(PG_ERR_INVALID, "batch outside 1..2^24");
    if (!(cfg.weight_clip > 0.f)) return c->fail(PG_ERR_INVALID, "weight_clip must be positive");
    c->free_mlp();
    auto& m = c->mlp;
    m.cfg = cfg;
    const size_t T = (size_t)1 << cfg.log2_table;
    m.n_grid = (size_t)kHashLevels * T * 2;
    PG_TRY(c, dalloc(&m.params, kParams), "alloc network");
    PG_TRY(c, dalloc(&m.pm, kParams), "alloc network");
    PG_TRY(c, dalloc(&m.pv, kParams), "alloc network");
    PG_TRY(c, dalloc(&m.mirror, kMirror), "alloc network");
    PG_TRY(c, dalloc(&m.grid, m.n_grid), "alloc hash grid");
    PG_TRY(c, dalloc(&m.gm, m.n_grid), "alloc hash grid");
    PG_TRY(c, dalloc(&m.gv, m.n_grid), "alloc hash grid");
    // view first: the gradient accumulator layout depends on the level resolutions
    PG_TRY(c, dalloc(&m.grid_bf, m.n_grid / 2), "alloc hash grid");
    PG_TRY(c, dalloc(&m.status, 1), "alloc status");
    PG_TRY(c, dalloc(&m.sum_part, kSumParts), "alloc status");
    // initialisation: Xavier-uniform weights over the used inputs/outputs, zero
    // biases and padding, hash features uniform in +-1e-4 (counter hash of the seed)
    std::vector<float> hp(kParams, 0.f);
    auto uni = [&](uint32_t stream, uint32_t i) { return keyed_rand(path_key(cfg.seed, stream, i), 0) * 2.f - 1.f; };
    auto fill = [&](int off, int rows, int cols, int used_rows, int used_cols, uint32_t stream) {
        float a = std::sqrt(6.f / (float)(used_rows + used_cols));
        for (int r = 0; r < used_rows; ++r)
            for (int q = 0; q < used_cols; ++q) hp[off + r * cols + q] = a * uni(stream, (uint32_t)(r * cols + q));
        (void)rows;
    };
    fill(kOffW1, kHidden, kInputs, kHidden, kInputsUsed, 11);
    fill(kOffW2, kHidden, kHidden, kHidden, kHidden, 12);
    fill(kOffW3, kHidden, kHidden, kHidden, kHidden, 13);
    fill(kOffW4, kOutputs, kHidden, kOutputsUsed, kHidden, 14);
    std::vector<float> hg(m.n_grid);
    for (size_t i = 0; i < m.n_grid; ++i) hg[i] = 1e-4f * uni(15, (uint32_t)i);
    PG_TRY(c, hipMemcpy(m.params, hp.data(), kParams * 4, hipMemcpyHostToDevice), "init network");
    PG_TRY(c, hipMemcpy(m.grid, hg.data(), m.n_grid * 4, hipMemcpyHostToDevice), "init hash grid");
    PG_TRY(c, hipMemset(m.pm, 0, kParams * 4), "init moments");
    PG_TRY(c, hipMemset(m.pv, 0, kParams * 4), "init moments");
    PG_TRY(c, hipMemset(m.gm, 0, m.n_grid * 4), "init moments");
    PG_TRY(c, hipMemset(m.gv, 0, m.n_grid * 4), "init moments");
    PG_TRY(c, hipMemset(m.mirror, 0, kMirror * 2), "init mirror");
    PG_TRY(c, hipMemset(m.status, 0, 4), "init status");
    PG_TRY(c, launch_mlp_mirror(m.params, m.mirror, c->stream), "mirror");
    PG_TRY(c, launch_grid_mirror(m.grid, m.grid_bf, m.n_grid, c->stream), "mirror");
    PG_TRY(c, hipStreamSynchronize(c->stream), "init");
    MlpView& v = m.view;
    v.wmirror = m.mirror;
    v.params = m.params;
    v.grid = m.grid_bf;
    mlp_view_frame(v, cfg.log2_table, cfg.res_min, cfg.res_max, c->host.centre, c->host.radius);
    uint64_t goff = 0;
    for (int l = 0; l < kHashLevels; ++l) {
        uint64_t ents = v.dense[l] ? (uint64_t)(v.res[l] + 1) * (v.res[l] + 1) * (v.res[l] + 1) : (uint64_t)T;
        v.gents[l] = (uint32_t)ents;
        uint32_t copies = 1;
        while (copies < kGradCopiesMax && (uint64_t)copies * 2u * ents <= kGradHitsPerCopy) copies *= 2;
        v.gcopies[l] = copies;
        v.goff[l] = goff;
        goff += ents * 2u * v.gcopies[l];
    }
    m.n_ggrad = goff;
    PG_TRY(c, dalloc(&m.ggrad, m.n_ggrad), "alloc hash-grid gradients");
    PG_TRY(c, hipMemset(m.ggrad, 0, m.n_ggrad * 4), "init gradients");
    m.ready = true;
    m.online_passes = m.record_passes = 0;
    // the first online record pass starts a fresh window: records left by earlier
    // passes (another guide's record pass, pg_records_upload) never reach training
    m.window_passes = ~0u;
    return PG_OK;
}

pg_status pg_mlp_train(pg_ctx* c, uint32_t steps, float* losses) {
    PG_CHECK_CTX(c);
    auto& m = c->mlp;
    if (!m.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    if (c->n_records == 0) return c->fail(PG_ERR_STATE, "no records to train on");
    if (c->n_aux != c->n_records)
        return c->fail(PG_ERR_STATE, "records carry no training data (only record passes keep it)");
    if (c->n_records > 0xffffffffull) return c->fail(PG_ERR_OVERFLOW, "more than 2^32 records");
    if (steps == 0) return PG_OK;
    const uint32_t n = (uint32_t)c->n_records;
    const uint32_t B = m.cfg.batch;
    pg_status st = mlp_buffers(c, B, steps);
    if (st != PG_OK) return st;
    if (m.w_scale_gen != c->rec_gen) {
        // mean sample weight over every record (deterministic one-block sum)
        PG_TRY(c, launch_sum(reinterpret_cast<const float*>(c->rec_aux) + 3, n, 8, m.sum_part, m.scal, c->stream), "weight sum");
        float sum = 0.f;
        PG_TRY(c, hipMemcpyAsync(&sum, m.scal, 4, hipMemcpyDeviceToHost, c->stream), "weight sum");
        PG_TRY(c, hipStreamSynchronize(c->stream), "weight sum");
        double mean = (double)sum / (double)n;
        m.w_scale = (mean > 0.0 && std::isfinite(mean)) ? (float)(1.0 / mean) : 1.f;
        m.w_scale_gen = c->rec_gen;
    }
    static const uint32_t kPrimes[5] = {2654435761u, 2246822519u, 3266489917u, 668265263u, 374761393u};
    hipEvent_t e0 = c->ev(), e1 = c->ev();
    (void)hipEventRecord(e0, c->stream);
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t t = m.step + 1;
        AdamParams ap{m.cfg.lr, m.cfg.beta1, m.cfg.beta2, m.cfg.eps,
                      (float)(1.0 - std::pow((double)m.cfg.beta1, (double)t)),
                      (float)(1.0 - std::pow((double)m.cfg.beta2, (double)t))};
        uint32_t offset = hash32(m.cfg.seed ^ hash32(t * 0x9E3779B9u)) % n;
        uint32_t stride = 1;
        for (int q = 0; q < 5; ++q) {
            uint32_t p = kPrimes[(t + q) % 5];
            if (p % n != 0) { stride = p % n; break; }
        }
        PG_TRY(c, launch_mlp_gather(c->records, c->rec_aux, n, offset, stride, B, m.w_scale, m.cfg.weight_clip,
                                    m.batch_raw, c->stream), "mlp gather");
        PG_TRY(c, launch_mlp_sort_batch(m.view, m.batch_raw, B, m.sort, m.batch, c->stream), "mlp batch order");
        uint32_t groups = 0;
        st = mlp_backward(c, B, m.cfg.nan_step == t ? 1 : 0, &groups);
        if (st != PG_OK) return st;
        PG_TRY(c, launch_mlp_adam(m.params, m.pm, m.pv, m.mirror, m.partial, groups, ap, 1, nullptr, m.status,
                                  c->stream), "mlp adam");
        PG_TRY(c, launch_grid_adam(m.view, m.grid, m.grid_bf, m.gm, m.gv, m.ggrad, m.n_grid, ap, 1, nullptr, m.status,
                                   c->stream), "grid adam");
        PG_TRY(c, launch_sum(m.loss, B, 1, m.sum_part, m.scal + k, c->stream), "loss");
        m.step = t;
    }
    (void)hipEventRecord(e1, c->stream);
    std::vector<float> hl(steps);
    uint32_t status = 0;
    PG_TRY(c, hipMemcpyAsync(hl.data(), m.scal, steps * 4, hipMemcpyDeviceToHost, c->stream), "read losses");
    PG_TRY(c, hipMemcpyAsync(&status, m.status, 4, hipMemcpyDeviceToHost, c->stream), "read status");
    PG_TRY(c, hipStreamSynchronize(c->stream), "train");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c->pool_used = 0;
    m.train_ms = ms / (float)steps;  // per Adam step
    m.last_loss = hl[steps - 1];
    for (float l : hl)
        if (!std::isfinite(l)) status |= 4u;
    if (status && !m.diverged) {
        m.diverged = true;
        m.diverge_reason = PG_MLP_DIVERGED_NONFINITE;
    }
    // rising loss: the loss smoothed over steps (weight 1/4 on the newest) staying
    // more than diverge_rise nats above its best value for diverge_patience steps
    const float rise = m.cfg.diverge_rise == 0.f ? 2.f : m.cfg.diverge_rise;
    // (16 steps by default: online training takes several steps on each record pass,
    // and one pass's records can raise the loss for all of its steps)
    const uint32_t patience = m.cfg.diverge_patience ? m.cfg.diverge_patience : 16u;
    for (uint32_t k = 0; k < steps && !m.diverged && rise > 0.f; ++k) {
        float l = hl[k];
        const uint32_t t = m.step - steps + 1 + k;  // Adam step that produced hl[k]
        if (m.cfg.rise_test_step && t >= m.cfg.rise_test_step) l += 10.f * (float)(t - m.cfg.rise_test_step + 1);
        m.loss_ema = m.loss_seen == 0 ? l : 0.75f * m.loss_ema + 0.25f * l;
        ++m.loss_seen;
        if (m.loss_seen <= patience || m.loss_ema < m.loss_best) {
            if (m.loss_seen == 1 || m.loss_ema < m.loss_best) m.loss_best = m.loss_ema;
            m.rise_run = 0;
            continue;
        }
        m.rise_run = m.loss_ema > m.loss_best + rise ? m.rise_run + 1 : 0;
        if (m.rise_run >= patience) {
            m.diverged = true;
            m.diverge_reason = PG_MLP_DIVERGED_RISING_LOSS;
        }
    }
    if (losses) std::memcpy(losses, hl.data(), steps * 4);
    return PG_OK;
}

pg_status pg_mlp_render_online(pg_ctx* c, const pg_render_params* params, uint32_t n_spp,
                                const pg_mlp_schedule* schedule) {
    PG_CHECK_CTX(c);
    if (!c->mlp.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    pg_mlp_schedule sch{1, 1, 1, 0, 0, 1, {0, 0}};
    if (schedule) sch = *schedule;
    if (sch.train_every == 0) sch.train_every = 1;
    if (sch.window == 0) sch.window = 1;
    if (sch.record_samples == 0) sch.record_samples = 1;
    if (sch.record_samples > 1 && sch.record_samples >= sch.train_every)
        return c->fail(PG_ERR_INVALID, "record_samples must be below train_every");
    pg_render_params p = params ? *params : pg_render_params_default();
    p.integrator = PG_INTEGRATOR_GUIDED_MLP;
    auto& m = c->mlp;
    for (uint32_t s = 0; s < n_spp;) {
        const bool record = (m.online_passes % sch.train_every) == 0;
        if (!record) {
            // the guided passes up to the next record pass go in one call, so that they
            // share wavefront passes (samples_per_pass)
            uint32_t k = 1;
            while (s + k < n_spp && ((m.online_passes + k) % sch.train_every) != 0) ++k;
            m.online_passes += k;
            s += k;
            p.record = 0;
            pg_status st = pg_render_spp(c, &p, k);
            if (st != PG_OK) return st;
            continue;
        }
        // a record pass of record_samples samples (fewer when the call ends first)
        const uint32_t rs = std::min(sch.record_samples, n_spp - s);
        s += rs;
        m.online_passes += rs;
        p.record = 1;
        if (m.window_passes >= sch.window) {
            c->n_records = 0;
            c->n_aux = 0;
            c->rec_gen++;
            c->cache_ready = false;
            m.window_passes = 0;
        }
        pg_status st = pg_render_spp(c, &p, rs);
        if (st != PG_OK) return st;
        m.window_passes++;
        m.record_passes++;
        uint32_t steps = m.record_passes <= sch.warmup_passes ? sch.warmup_steps : sch.steps;
        if (steps && !m.diverged && c->n_records > 0) {
            st = pg_mlp_train(c, steps, nullptr);
            if (st != PG_OK) return st;
        }
    }
    return PG_OK;
}

pg_status pg_mlp_info_read(pg_ctx* c, pg_mlp_info* info) {
    PG_CHECK_CTX(c);
    if (!info) return c->fail(PG_ERR_INVALID, "null info");
    const auto& m = c->mlp;
    std::memset(info, 0, sizeof(*info));
    info->ready = m.ready ? 1u : 0u;
    info->diverged = m.diverged ? 1u : 0u;
    info->step = m.step;
    info->n_params = kParams;
    info->n_grid_params = m.n_grid;
    info->last_loss = m.last_loss;
    info->w_scale = m.w_scale;
    info->chunks = m.chunks;
    info->train_ms = m.train_ms;
    info->fallbacks = m.fallbacks;
    info->diverge_reason = m.diverge_reason;
    return PG_OK;
}

pg_status pg_mlp_get_params(pg_ctx* c, float* params, float* grid) {
    PG_CHECK_CTX(c);
    const auto& m = c->mlp;
    if (!m.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    if (params) PG_TRY(c, hipMemcpy(params, m.params, kParams * 4, hipMemcpyDeviceToHost), "read network");
    if (grid) PG_TRY(c, hipMemcpy(grid, m.grid, m.n_grid * 4, hipMemcpyDeviceToHost), "read hash grid");
    return PG_OK;
}

pg_status pg_mlp_set_params(pg_ctx* c, const float* params, const float* grid) {
    PG_CHECK_CTX(c);
    auto& m = c->mlp;
    if (!m.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    if (params) {
        PG_TRY(c, hipMemcpy(m.params, params, kParams * 4, hipMemcpyHostToDevice), "write network");
        PG_TRY(c, hipMemset(m.pm, 0, kParams * 4), "reset moments");
        PG_TRY(c, hipMemset(m.pv, 0, kParams * 4), "reset moments");
        PG_TRY(c, launch_mlp_mirror(m.params, m.mirror, c->stream), "mirror");
    }
    if (grid) {
        PG_TRY(c, hipMemcpy(m.grid, grid, m.n_grid * 4, hipMemcpyHostToDevice), "write hash grid");
        PG_TRY(c, hipMemset(m.gm, 0, m.n_grid * 4), "reset moments");
        PG_TRY(c, hipMemset(m.gv, 0, m.n_grid * 4), "reset moments");
        PG_TRY(c, launch_grid_mirror(m.grid, m.grid_bf, m.n_grid, c->stream), "mirror");
    }
    PG_TRY(c, hipStreamSynchronize(c->stream), "set params");
    return PG_OK;
}

pg_status pg_mlp_levels(pg_ctx* c, uint32_t* res, uint32_t* dense, float* frame) {
    PG_CHECK_CTX(c);
    const auto& m = c->mlp;
    if (!m.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    for (int l = 0; l < kHashLevels; ++l) {
        if (res) res[l] = m.view.res[l];
        if (dense) dense[l] = m.view.dense[l];
    }
    if (frame) {
        frame[0] = m.view.lo[0]; frame[1] = m.view.lo[1]; frame[2] = m.view.lo[2];
        frame[3] = m.view.inv_ext;
    }
    return PG_OK;
}

pg_status pg_mlp_eval(pg_ctx* c, uint32_t n, const float* pos, const float* nrm, const float* wo, float* mix,
                      const float* wi_eval, float* pdf_eval, const float* u3, float* wi_out) {
    PG_CHECK_CTX(c);
    const auto& m = c->mlp;
    if (!m.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    if (n == 0) return PG_OK;
    if (!pos || !nrm || !wo) return c->fail(PG_ERR_INVALID, "null inputs");
    if ((wi_eval != nullptr) != (pdf_eval != nullptr) || (u3 != nullptr) != (wi_out != nullptr))
        return c->fail(PG_ERR_INVALID, "wi_eval/pdf_eval and u3/wi_out come in pairs");
    auto pack4 = [n](const float* p) {
        std::vector<float4> v(n);
        for (uint32_t i = 0; i < n; ++i) v[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], 0.f);
        return v;
    };
    std::vector<float4> hx = pack4(pos), hn = pack4(nrm), hw = pack4(wo);
    float4 *dx = nullptr, *dn = nullptr, *dw = nullptr, *dm = nullptr, *de = nullptr, *du = nullptr, *dwo = nullptr;
    float* dpe = nullptr;
    pg_status st = PG_OK;
    do {
        if ((st = c->hip(dalloc(&dx, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dn, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dw, n), "alloc")) != PG_OK) break;
        if ((st = c->hip(dalloc(&dm, (size_t)n * kMixF4), "alloc")) != PG_OK) break;
        hipMemcpy(dx, hx.data(), (size_t)n * 16, hipMemcpyHostToDevice);
        hipMemcpy(dn, hn.data(), (size_t)n * 16, hipMemcpyHostToDevice);
        if ((st = c->hip(hipMemcpy(dw, hw.data(), (size_t)n * 16, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        if ((st = c->hip(launch_mlp_eval(m.view, dx, dn, dw, n, dm, c->cu_count, c->stream), "mlp eval")) != PG_OK)
            break;
        if (wi_eval) {
            std::vector<float4> he = pack4(wi_eval);
            if ((st = c->hip(dalloc(&de, n), "alloc")) != PG_OK) break;
            if ((st = c->hip(dalloc(&dpe, n), "alloc")) != PG_OK) break;
            if ((st = c->hip(hipMemcpy(de, he.data(), (size_t)n * 16, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        }
        if (u3) {
            std::vector<float4> hu = pack4(u3);
            if ((st = c->hip(dalloc(&du, n), "alloc")) != PG_OK) break;
            if ((st = c->hip(dalloc(&dwo, n), "alloc")) != PG_OK) break;
            if ((st = c->hip(hipMemcpy(du, hu.data(), (size_t)n * 16, hipMemcpyHostToDevice), "copy")) != PG_OK) break;
        }
        if (de || du)
            if ((st = c->hip(launch_mix_query(dm, n, de, du, dpe, dwo, c->stream), "mixture query")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        if (mix && (st = c->hip(hipMemcpy(mix, dm, (size_t)n * kMixF4 * 16, hipMemcpyDeviceToHost), "copy")) != PG_OK)
            break;
        if (pdf_eval && (st = c->hip(hipMemcpy(pdf_eval, dpe, (size_t)n * 4, hipMemcpyDeviceToHost), "copy")) != PG_OK)
            break;
        if (wi_out && (st = c->hip(hipMemcpy(wi_out, dwo, (size_t)n * 16, hipMemcpyDeviceToHost), "copy")) != PG_OK)
            break;
    } while (0);
    dfree(dx); dfree(dn); dfree(dw); dfree(dm); dfree(de); dfree(du); dfree(dwo); dfree(dpe);
    return st;
}

pg_status pg_mlp_gradient(pg_ctx* c, uint32_t n, const float* samples, float* loss, float* grad_params,
                          float* grad_grid) {
    PG_CHECK_CTX(c);
    auto& m = c->mlp;
    if (!m.ready) return c->fail(PG_ERR_STATE, "pg_mlp_init first");
    if (n == 0 || !samples) return c->fail(PG_ERR_INVALID, "no samples");
    if (n > (1u << 24)) return c->fail(PG_ERR_INVALID, "more than 2^24 samples");
    pg_status st = mlp_buffers(c, n, 1);
    if (st != PG_OK) return st;
    PG_TRY(c, hipMemcpy(m.batch, samples, (size_t)n * sizeof(TrainSample), hipMemcpyHostToDevice), "upload samples");
    float* dgp = nullptr;
    float* dgg = nullptr;
    do {
        if ((st = c->hip(dalloc(&dgp, kParams), "alloc")) != PG_OK) break;
        if (grad_grid && (st = c->hip(dalloc(&dgg, m.n_grid), "alloc")) != PG_OK) break;
        uint32_t groups = 0;
        if ((st = mlp_backward(c, n, 0, &groups)) != PG_OK) break;
        AdamParams ap{m.cfg.lr, m.cfg.beta1, m.cfg.beta2, m.cfg.eps, 1.f, 1.f};
        if ((st = c->hip(launch_mlp_adam(m.params, m.pm, m.pv, m.mirror, m.partial, groups, ap, 0, dgp, m.status,
                                         c->stream), "gradient")) != PG_OK) break;
        if ((st = c->hip(launch_grid_adam(m.view, m.grid, m.grid_bf, m.gm, m.gv, m.ggrad, m.n_grid, ap, 0, dgg, m.status,
                                          c->stream), "gradient")) != PG_OK) break;
        if ((st = c->hip(launch_sum(m.loss, n, 1, m.sum_part, m.scal, c->stream), "loss")) != PG_OK) break;
        if ((st = c->hip(hipStreamSynchronize(c->stream), "sync")) != PG_OK) break;
        if (loss && (st = c->hip(hipMemcpy(loss, m.scal, 4, hipMemcpyDeviceToHost), "copy")) != PG_OK) break;
        if (grad_params && (st = c->hip(hipMemcpy(grad_params, dgp, kParams * 4, hipMemcpyDeviceToHost), "copy")) != PG_OK)
            break;
        if (grad_grid && (st = c->hip(hipMemcpy(grad_grid, dgg, m.n_grid * 4, hipMemcpyDeviceToHost), "copy")) != PG_OK)
            break;
    } while (0);
    dfree(dgp);
    dfree(dgg);
    return st;
}

}  // extern "C"
