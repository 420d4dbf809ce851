/*
 * pg_tracer.h — C ABI of the MI355X path-tracing hot path.
 *
 * The reference (IsaGeriler/PathGuidingDissertation) has no FFI: its hot path is a
 * set of C++ classes called directly (Scene, BVHNode, BSDF, Light, RayTracer). Each
 * entry point below names the reference interface it replaces (file:line, relative
 * to PathGuidingDissertation/). Host C++ classes mirroring those interfaces
 * (pathguidingdissertation_amd/include/pgd/*.hpp) flatten into these calls.
 *
 * Conventions: every function returns pg_status (0 = OK, negative = error) and
 * never throws; pg_last_error() gives the message. Caller-owned host arrays are
 * copied in/out; the context owns device memory. One context per device and
 * host thread; calls are ordered on the context's HIP stream and return after
 * the work completes. Results are deterministic for a given (seed, params).
 */
#ifndef PG_TRACER_H
#define PG_TRACER_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pg_status;
#define PG_OK 0
#define PG_ERR_INVALID (-1)   /* bad argument or size                          */
#define PG_ERR_NOMEM (-2)     /* device or host allocation failed              */
#define PG_ERR_DEVICE (-3)    /* HIP runtime error / no device                 */
#define PG_ERR_OVERFLOW (-4)  /* queue, stack or capacity exceeded             */
#define PG_ERR_STATE (-5)     /* call out of order (e.g. trace before build)   */

typedef struct pg_ctx pg_ctx;

/* ------------------------------------------------------------------ data formats */

/* Material (Materials.h:121-897; parameters as SceneLoader.h:104-181 reads them).
 * albedo is a constant colour unless pg_scene_textures gives the material an
 * albedo texture (the reference's Texture* albedo, Imaging.h:82-108).
 * PG_BSDF_LAYERED added to a base type is LayeredBSDF over that base (Materials.h:
 * 855-897: every call forwards to the base, isTwoSided is true).                */
enum { PG_BSDF_DIFFUSE = 0, PG_BSDF_MIRROR = 1, PG_BSDF_CONDUCTOR = 2, PG_BSDF_GLASS = 3,
       PG_BSDF_DIELECTRIC = 4, PG_BSDF_ORENNAYAR = 5, PG_BSDF_PLASTIC = 6, PG_BSDF_LAYERED = 8 };
typedef struct {
    float type;          /* PG_BSDF_* stored as float so the record is 16 floats   */
    float albedo[3];
    float emission[3];   /* > 0 luminance makes every triangle an AreaLight        */
    float eta[3], k[3];  /* conductor                                              */
    float int_ior, ext_ior; /* glass, dielectric, plastic                          */
    float roughness;     /* alpha = 1.62142 * sqrt(roughness) (Materials.h:249, 738);
                            Oren-Nayar: sigma itself (Materials.h:649-652)         */
} pg_material;

/* Pinhole camera as the scene loader builds it (SceneLoader.h:226-244). */
typedef struct {
    float from[3], to[3], up[3];
    float fov_deg;
    float width, height;
} pg_camera_desc;

/* Ray for the trace entry points: origin, direction (Geometry.h:22-39). */
typedef struct { float o[3]; float d[3]; } pg_ray;

/* Closest hit (Geometry.h:247-253): ID = UINT32_MAX, t = FLT_MAX on a miss. */
typedef struct { uint32_t id; float t, alpha, beta, gamma; } pg_hit;

/* Path-vertex record (Renderer.h:42-47), 48 bytes in HBM. */
typedef struct { float pos[3]; float normal[3]; float wi[3]; float li[3]; } pg_path_vertex;

/* PG_INTEGRATOR_GUIDED_PSS: the guiding the reference outlines (Renderer.h:398-414):
 * at each vertex the k nearest records' directions are projected into the BSDF's
 * primary sample space by BSDF::invert, binned by flux over a 4x4 (u, v) grid, and
 * the grid is sampled and mapped back through the BSDF's warp (one-sample MIS with
 * plain BSDF sampling). Guides Diffuse, Oren-Nayar and rough Conductor vertices;
 * needs pg_cache_build like the quadtree guide. knn_k 16: exact 16 nearest records;
 * 0: approximate neighbours (cheaper).                                           */
enum { PG_INTEGRATOR_PATH = 0, PG_INTEGRATOR_RECORD = 1, PG_INTEGRATOR_GUIDED_QT = 2,
       PG_INTEGRATOR_GUIDED_MLP = 3, PG_INTEGRATOR_GUIDED_PSS = 4 };

/* Render parameters: the reference's hard-coded constants become fields whose
 * defaults (pg_render_params_default) equal the reference.                    */
typedef struct {
    uint32_t seed;
    uint32_t max_depth;     /* Renderer.h:568 -> 8                              */
    uint32_t rr_depth;      /* Renderer.h:561 -> Russian roulette when depth > 3 */
    int32_t integrator;     /* PG_INTEGRATOR_*                                  */
    uint32_t record;        /* RECORD integrator: keep PathVertex records       */
    float guide_prob;       /* guided integrators: P(sample guide), 0.5; the
                               learned guide samples with its own per-point
                               probability, capped at guide_prob (0 disables it) */
    uint32_t knn_k;         /* guided QT / PSS: neighbours per lookup (16); PSS
                               with 0 takes the 16 records nearest the centroid of
                               the point-BVH leaf that holds the query's Morton key
                               (lists built with the cache) instead of a search   */
    uint32_t record_nee;    /* RECORD: 0 records Li = G (Renderer.h:390, the
                               reference); 1 records Li with NEE and MIS-weighted
                               emission (the pixelColour recursion, :391)       */
    uint32_t trace_blocks_per_cu;  /* traversal grid (0: as many as LDS allows)   */
    uint32_t shade_blocks_per_cu;  /* shade grid, blocks of 256 lanes (0: 3)     */
    uint32_t sort_rays;     /* 1: trace secondary and shadow rays in (direction
                               octant, origin Morton) order; results unchanged  */
    uint32_t dyn_fetch;     /* 1 (default): persistent lanes claim new rays from
                               their wave's share of the queue when the wave runs
                               below 48 busy lanes; 0: one ray per lane and grid
                               stride. Results unchanged                         */
    uint32_t env_visibility; /* environment-light shadow rays: 0 aim at
                               sceneCentre + wi * sceneRadius as Renderer.h:331-332
                               does (default, the reference's images); 1 trace
                               along wi out of the bounding sphere. With 0 a point
                               away from the centre can test the wrong segment
                               (behind a small window: direct sun lost), and then
                               guided and unguided MIS estimators converge to
                               different images; 1 makes them agree.            */
    uint32_t samples_per_pass; /* samples per pixel traced together in one
                               wavefront pass (the deep bounces' few rays then
                               share launches); 0 = automatic: the call's SPP
                               split evenly over passes of at most 2^26 paths
                               (about 14 GB of path state, twice that with a
                               guide; record passes at most
                               min(2^22, 2^22 x 13 / (max_depth + 1)) paths:
                               their vertex stacks take (max_depth + 1) x 128 B
                               per path, ~7 GB).
                               Images and records are identical for every value. */
    uint32_t stack_lds;     /* traversal stack entries per lane kept in LDS (16, 32,
                               48 or 64), deeper entries in global memory: fewer
                               take less LDS per block and fit more blocks per CU.
                               0 = automatic (16). Results unchanged.            */
    uint32_t guide_depth;   /* guided integrators: guide the bounces below this
                               depth and sample the BSDF alone deeper (0 = every
                               bounce); single-view and batch renders            */
} pg_render_params;

typedef struct {
    uint64_t closest_rays, shadow_rays, paths, records;
    double kernel_ms_extend, kernel_ms_shade, kernel_ms_shadow, kernel_ms_total;
    uint64_t extend_launches, shadow_launches;
    uint64_t nodes_visited, tris_tested;       /* filled when counting is enabled */
    uint64_t shadow_nodes_visited, shadow_tris_tested;
    double kernel_ms_guide;  /* guide queries per bounce: kNN / query order (quadtree,
                                PSS guides) or network inference (learned guide)     */
} pg_stats;

/* ------------------------------------------------------------------ context */
pg_status pg_create(pg_ctx** ctx, int device);          /* RayTracer::init Renderer.h:251-263 */
pg_status pg_destroy(pg_ctx* ctx);
const char* pg_last_error(const pg_ctx* ctx);            /* NULL ctx: last global error        */
pg_status pg_device_info(pg_ctx* ctx, char* name, size_t name_len, int32_t* cu_count,
                         double* peak_hbm_gbs);

/* ------------------------------------------------------------------ scene */
/* Scene::init + Scene::build (Scene.h:84-137): copies triangles (v0,v1,v2 positions
 * and per-vertex shading normals, 9 floats each), per-triangle material index,
 * the material table and an optional lat-long RGB32F environment map (y-up,
 * Lights.h:128-228; NULL => black BackgroundColour, Lights.h:75-126). The host
 * builds the binned-SAH BVH in the reference's topology (Geometry.h:293-446) and
 * uploads 32-byte nodes, leaf-ordered triangles, shading attributes, the light
 * list and the environment alias table. At most 2^26 triangles (PG_ERR_INVALID).*/
pg_status pg_scene_upload(pg_ctx* ctx, const float* positions, const float* normals,
                          const uint32_t* material_index, uint32_t n_tris,
                          const pg_material* materials, uint32_t n_materials,
                          const float* env_rgb, int32_t env_width, int32_t env_height,
                          const pg_camera_desc* camera);
/* Albedo textures of the uploaded scene (Texture::sample, Imaging.h:82-108, read at
 * the hit's interpolated uv, Geometry.h:132-137): uv holds n_tris*6 floats (u0 v0 u1
 * v1 u2 v2 per triangle, upload order); material_texture one index per material
 * (-1 keeps the constant albedo); texels the RGB of every texture concatenated,
 * texture i starting at texel texel_offset[i], tex_width[i] x tex_height[i]. A new
 * pg_scene_upload drops them. Needs an uploaded scene (PG_ERR_STATE).            */
pg_status pg_scene_textures(pg_ctx* ctx, const float* uv, const int32_t* material_texture, uint32_t n_materials,
                            const float* texels, const uint64_t* texel_offset, const int32_t* tex_width,
                            const int32_t* tex_height, uint32_t n_textures);
/* Where pg_scene_upload builds the BVH: PG_BVH_HOST (default; threads of the host)
 * or PG_BVH_GPU (the device, level by level). Both give the reference's topology,
 * node numbering and triangle order (Geometry.h:274-446) bit for bit.           */
enum { PG_BVH_HOST = 0, PG_BVH_GPU = 1 };
pg_status pg_set_bvh_builder(pg_ctx* ctx, int32_t builder);
/* BVHNode::build (Geometry.h:419-446) on the GPU for n_tris triangles (9 floats of
 * positions each): nodes_out receives the node count x 8 floats (min.xyz, leftFirst
 * bits, max.xyz, used bits; at most 2 n_tris - 1 nodes), tri_index_out the
 * triangleIndexes order, build_ms the GPU time. Same output as pg_bvh_build_host.  */
pg_status pg_bvh_build_gpu(pg_ctx* ctx, const float* positions, uint32_t n_tris, uint32_t* n_nodes, float* nodes_out,
                           int32_t* tri_index_out, double* build_ms);
/* BVH statistics: node count, maximum leaf depth, build milliseconds. */
pg_status pg_bvh_info(pg_ctx* ctx, uint32_t* n_nodes, uint32_t* max_depth, double* build_ms);
/* Export the topology as built (8 floats per node: min.xyz, leftFirst bits,
 * max.xyz, used bits) and the triangle order (BVHNode::triangleIndexes).        */
pg_status pg_bvh_export(pg_ctx* ctx, float* nodes, int32_t* tri_index);

/* ------------------------------------------------------------------ queries (tests) */
/* BVHNode::traverse (Geometry.h:448-512) for a batch of rays. */
pg_status pg_trace_closest(pg_ctx* ctx, const pg_ray* rays, uint32_t n, pg_hit* hits);
/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): pg_trace_closest without any allocation. Rays and hits stage through
 * one fixed device array per GPU (64 B per ray, at most 2^18 = 262,144 rays per call,
 * 16 MiB, held until the process ends); more rays are PG_ERR_INVALID. Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (several contexts or
 * host threads) race and must be serialised by the caller.                       */
pg_status pg_trace_closest_noalloc(pg_ctx* ctx, const pg_ray* rays, uint32_t n, pg_hit* hits);
/* BVHNode::traverseVisible (Geometry.h:514-567): visible[i] = 1 when no hit in
 * (1e-4, max_t[i]).                                                            */
pg_status pg_trace_shadow(pg_ctx* ctx, const pg_ray* rays, const float* max_t, uint32_t n,
                          uint8_t* visible);
/* Camera::generateRay (Scene.h:45-57) for film positions (x, y). */
pg_status pg_generate_rays(pg_ctx* ctx, const float* pxy, uint32_t n, pg_ray* rays);
/* Counter RNG values for (seed, sample, pixel, dim) — parity aid. */
pg_status pg_rng_values(pg_ctx* ctx, uint32_t seed, uint32_t sample, const uint32_t* pixel,
                        const uint32_t* dim, uint32_t n, float* out);

/* BSDF::sample / evaluate / PDF / invert / isPureSpecular / isTwoSided (Materials.h:
 * 121-897) of one material on the device, for n shading points with given random
 * numbers (the reference's test sampler, main.cpp:13-72) — parity aid. sd: 12 floats
 * per query (x, wo, shading normal, geometric normal; frame from the shading normal),
 * u: 3, wi_eval: 3. out: 15 floats per query: sampled wi[3], f[3], pdf, evaluate(
 * wi_eval)[3], PDF(wi_eval), invert(sampled wi) u, v (0, 0 where the BSDF has no
 * invert), isPureSpecular, isTwoSided. At most 2^24 queries.                     */
pg_status pg_bsdf_query(pg_ctx* ctx, const pg_material* material, uint32_t n, const float* sd, const float* u,
                        const float* wi_eval, float* out);

/* ------------------------------------------------------------------ rendering */
pg_render_params pg_render_params_default(void);
/* RayTracer::clear (Renderer.h:265-267): zero the film and SPP count. */
pg_status pg_film_clear(pg_ctx* ctx);
/* RayTracer::render (Renderer.h:648-747) n_spp times: every sample of every pixel
 * goes through the wavefront pipeline (ray-gen, extend, shade, shadow) and is
 * splatted with the reference's box filter (Imaging.h:186-217). Samples are traced
 * params->samples_per_pass at a time; the film (and a record pass's records) equal
 * n_spp calls of one sample each bit for bit.                                    */
pg_status pg_render_spp(pg_ctx* ctx, const pg_render_params* params, uint32_t n_spp);
/* Launch-shape tuner for pg_render_spp (no reference counterpart): times passes of
 * params->samples_per_pass samples (0: 1) — a warm-up pass, then two timed passes
 * (one from 8 samples per pass) — over candidate LDS stack sizes, traversal and shade
 * grids, persistent lanes and ray sorting (coordinate descent) and returns params
 * with the fastest shape; clears the film.                                      */
pg_status pg_render_tune(pg_ctx* ctx, const pg_render_params* params, pg_render_params* tuned,
                         double* ms_per_spp);
/* Film::film / Film::SPP (Imaging.h:178-272): accumulated RGB sums (W*H*3) and SPP. */
pg_status pg_film_read(pg_ctx* ctx, float* rgb_sum, uint32_t* spp);
/* Film::tonemap (Imaging.h:220-244) of the whole film on the device: mean over the
 * film's SPP, exposure, extended Reinhard with white point lmax_white, gamma 1/2.2,
 * 8-bit RGB (width*height*3 bytes) — the canvas RayTracer::savePNG writes. The
 * reference calls it with exposure 1 and LmaxWhite 10 (Renderer.h:705).          */
pg_status pg_film_tonemap(pg_ctx* ctx, float exposure, float lmax_white, uint8_t* rgb8);
pg_status pg_stats_read(pg_ctx* ctx, pg_stats* stats);
pg_status pg_stats_reset(pg_ctx* ctx);
/* Instrumented traversal: count node visits / triangle tests per ray (slower). */
pg_status pg_set_counting(pg_ctx* ctx, int32_t enable);
/* Per-launch HIP-event timing of the render kernels (default on; fills the
 * kernel_ms_extend / shade / shadow statistics). Off leaves kernel_ms_total only. */
pg_status pg_set_timing(pg_ctx* ctx, int32_t enable);
/* Measurement aid (no reference counterpart): a plain device-to-device streaming
 * copy of `bytes`, timed with HIP events, as the HBM reference point for achieved
 * bandwidth and for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings.          */
pg_status pg_copy_benchmark(pg_ctx* ctx, uint64_t bytes, uint32_t iters, double* ms_per_copy);

/* ------------------------------------------------------------------ guiding cache */
/* Records gathered by PG_INTEGRATOR_RECORD passes with params.record != 0
 * (RayTracer::perThreadPathVertexRecords gathered after the join, Renderer.h:720-741),
 * in (pass, path, depth) order; 48-byte PathVertex records (Renderer.h:42-47).  */
pg_status pg_records_count(pg_ctx* ctx, uint64_t* n);
pg_status pg_records_read(pg_ctx* ctx, pg_path_vertex* out, uint64_t cap, uint64_t* n);
pg_status pg_records_clear(pg_ctx* ctx);
/* Replace the record set (tests, or records kept from an earlier run). */
pg_status pg_records_upload(pg_ctx* ctx, const pg_path_vertex* records, uint64_t n);
/* PointBVHNode::buildPointBVHNode (Renderer.h:189-211) on the GPU: Morton keys and a
 * radix sort, leaves of <= 8 records under an implicit balanced tree; plus one
 * directional quadtree per cell of `cell_size` consecutive sorted records (0: 4096),
 * refined where a quadrant holds > 1% of the cell's flux, depth <= 10, at most
 * `qt_nodes_per_cell` nodes (0: 1024).                                          */
pg_status pg_cache_build(pg_ctx* ctx, uint32_t cell_size, uint32_t qt_nodes_per_cell);
/* PointBVHNode::stats (Renderer.h:219-233): records, nodes, depth, cells, build ms. */
pg_status pg_cache_info(pg_ctx* ctx, uint64_t* n_records, uint32_t* n_nodes, uint32_t* depth,
                        uint32_t* n_cells, double* build_ms);
/* Export for validation (PointBVHNode::validate, Renderer.h:144-217): node boxes
 * (8 floats each), sorted record positions (xyz + original index bits) and the
 * quadtree pools (8 words per node: 4 flux floats, 4 child indices).            */
pg_status pg_cache_export(pg_ctx* ctx, float* nodes, float* sorted_pos, float* quadtrees);
/* k nearest records (k <= 16) by squared distance, ties to the lower record index;
 * no counterpart in the reference, which has no query on its point BVH.         */
pg_status pg_knn(pg_ctx* ctx, const float* query_xyz, uint32_t nq, uint32_t k, uint32_t* idx,
                 float* d2);
/* Guide distribution at points: a direction sampled with u3 and its solid-angle
 * pdf, and the pdf of wi_eval (the guiding design of Renderer.h:398-414).      */
pg_status pg_guide_qt_eval(pg_ctx* ctx, const float* pos, const float* u3, const float* wi_eval,
                           uint32_t n, float* wi_out, float* pdf_sample, float* pdf_eval);
/* The primary-sample-space guide (PG_INTEGRATOR_GUIDED_PSS with knn_k = 0) at n
 * shading points of one scene material: the records of the point's leaf list
 * inverted through BSDF::invert (Materials.h:163-171, 287-306, 673-681) into a 4x4
 * flux histogram; a direction sampled with u3 (bin by flux, point in the bin, the
 * BSDF's warp) and its solid-angle pdf, and the pdf of wi_eval. nrm is the shading
 * normal on wo's side. Zeros where the guide does not apply (specular or
 * non-invertible material, no record flux).                                      */
pg_status pg_guide_pss_eval(pg_ctx* ctx, uint32_t material, const float* pos, const float* nrm,
                            const float* wo, const float* u3, const float* wi_eval, uint32_t n,
                            float* wi_out, float* pdf_sample, float* pdf_eval);

/* ------------------------------------------------------------------ learned guide */
/* No counterpart in the reference, whose guiding stops at recording PathVertex data
 * (Renderer.h:398-414): a hash-grid + one-blob encoded 3 x 64 ReLU network maps
 * (x, n, wo) to an 8-lobe vMF mixture and a BSDF/guide selection probability,
 * trained with Adam on the records of PG_INTEGRATOR_RECORD passes (radiance-weighted
 * NLL of the one-sample mixture) and used by PG_INTEGRATOR_GUIDED_MLP.            */
typedef struct pg_mlp_config {
    uint32_t seed;            /* weight initialisation                              */
    uint32_t log2_table;      /* hash entries per level, log2 (12..22, default 19)  */
    float res_min, res_max;   /* coarsest / finest grid resolution (16, 1024)       */
    float lr, beta1, beta2, eps; /* Adam (1e-2, 0.9, 0.99, 1e-15)                   */
    uint32_t batch;           /* training samples per step (2^18)                   */
    float weight_clip;        /* clip of the mean-normalised sample weight (100)    */
    uint64_t act_budget;      /* bytes of activation rows per chunk (512 MiB)       */
    uint32_t nan_step;        /* test hook: poison the gradient at this step (0 off)*/
    float diverge_rise;       /* divergence on rising loss: the loss smoothed over
                                 steps (newest weighted 1/4) more than this many
                                 nats above its best value (0: 2; < 0: off) ...  */
    uint32_t diverge_patience; /* ... for this many consecutive steps (0: 16)    */
    uint32_t rise_test_step;  /* test hook: from this step on the detector sees the
                                 loss raised by 10 nats per step (0 off)         */
    uint32_t reserved[4];
} pg_mlp_config;

/* Why a learned guide was marked diverged (pg_mlp_info.diverge_reason). A diverged
 * guide is not trained further, and PG_INTEGRATOR_GUIDED_MLP renders with the
 * quadtree guide when pg_cache_build has built one (unguided otherwise).          */
enum { PG_MLP_DIVERGED_NONE = 0, PG_MLP_DIVERGED_NONFINITE = 1, PG_MLP_DIVERGED_RISING_LOSS = 2 };

typedef struct pg_mlp_info {
    uint32_t ready;           /* pg_mlp_init done                                   */
    uint32_t diverged;        /* a non-finite loss or gradient, or a rising loss, was
                                 seen: rendering falls back (diverge_reason)       */
    uint32_t step;            /* Adam steps taken                                   */
    uint32_t n_params;        /* network parameters (17648 incl. padding)           */
    uint64_t n_grid_params;   /* hash-table parameters (16 * 2^log2_table * 2)      */
    float last_loss;          /* mean loss of the last step's batch                 */
    float w_scale;            /* 1 / mean record weight                             */
    uint32_t chunks;          /* chunks per step under the activation budget        */
    float train_ms;           /* GPU time per Adam step of the last pg_mlp_train call */
    uint32_t fallbacks;       /* guided passes rendered with a fallback guide       */
    uint32_t diverge_reason;  /* PG_MLP_DIVERGED_*                                  */
    uint32_t reserved[4];
} pg_mlp_info;

pg_mlp_config pg_mlp_config_default(void);
/* Allocate and initialise the guide (needs an uploaded scene: its bounding cube
 * normalises positions; uploading another scene requires a new pg_mlp_init).
 * Resets the step count and the divergence flag.                                 */
pg_status pg_mlp_init(pg_ctx* ctx, const pg_mlp_config* cfg);
/* `steps` Adam steps on the current records (they must come from record passes,
 * which keep the training data). losses (nullable) receives each step's mean loss. */
pg_status pg_mlp_train(pg_ctx* ctx, uint32_t steps, float* losses);
/* Learned-guide rendering with online training. Every `train_every` SPP (the first
 * included) begin with a guided record pass of `record_samples` samples per pixel
 * (PG_INTEGRATOR_GUIDED_MLP with record = 1: the record integrator of
 * Renderer.h:367-529 sampling by one-sample MIS between the BSDF and the guide; its
 * samples go to the film and their vertices become training data, traced together as
 * one wavefront pass) followed by `steps` Adam steps on the records of the last
 * `window` record passes (older ones are dropped); the first `warmup_passes` record
 * passes train `warmup_steps` steps instead. The other SPP are guided path-tracing
 * passes (record = 0); those between two record passes are rendered by one call, so
 * they share wavefront passes (params->samples_per_pass). params->record_nee = 1 is
 * recommended. record_samples must be below train_every (PG_ERR_INVALID).          */
typedef struct pg_mlp_schedule {
    uint32_t train_every;    /* 1                                                  */
    uint32_t steps;          /* 1                                                  */
    uint32_t window;         /* 1                                                  */
    uint32_t warmup_passes;  /* 0                                                  */
    uint32_t warmup_steps;   /* 0                                                  */
    uint32_t record_samples; /* samples per pixel of a record pass; 0 = 1          */
    uint32_t reserved[2];
} pg_mlp_schedule;
pg_status pg_mlp_render_online(pg_ctx* ctx, const pg_render_params* params, uint32_t n_spp,
                               const pg_mlp_schedule* schedule);
pg_status pg_mlp_info_read(pg_ctx* ctx, pg_mlp_info* info);
/* FP32 master parameters (17648) and hash table (n_grid_params), either nullable. */
pg_status pg_mlp_get_params(pg_ctx* ctx, float* params, float* grid);
pg_status pg_mlp_set_params(pg_ctx* ctx, const float* params, const float* grid);
/* Per-level resolution and dense flag (16 each) and the position frame (lo.xyz, 1/extent). */
pg_status pg_mlp_levels(pg_ctx* ctx, uint32_t* res, uint32_t* dense, float* frame);
/* Network at n points (xyz triples): mixture (n x 52 floats, layout of 13 float4:
 * 8 x (mu, kappa), 8 lobe probabilities, 8 lobe probabilities x vMF normalisation,
 * (c, 0, 0, 0)), pdf of wi_eval, and a direction sampled with u3 (xyz + pdf).
 * Every output is nullable.                                                       */
pg_status pg_mlp_eval(pg_ctx* ctx, uint32_t n, const float* pos, const float* nrm, const float* wo,
                      float* mix, const float* wi_eval, float* pdf_eval, const float* u3,
                      float* wi_out);
/* Loss and gradients (no update) for n explicit samples of 16 floats each:
 * (x.xyz, weight), (n.xyz, BSDF pdf of wi), (wo.xyz, 0), (wi.xyz, 0); the loss is
 * the mean over the n samples. grad_params: 17648 floats; grad_grid: n_grid_params. */
pg_status pg_mlp_gradient(pg_ctx* ctx, uint32_t n, const float* samples, float* loss,
                          float* grad_params, float* grad_grid);

/* ------------------------------------------------------------------ batches of views */
/* No counterpart in the reference, which renders one camera per RayTracer::render
 * (Renderer.h:648-747): many views of one or more scenes rendered together. Every
 * pass runs all views' paths through the same persistent (grid-stride) wavefront
 * kernels in view-aligned chunks of at most queue_paths paths; with use_graph the
 * pass is captured once as a hipGraph and replayed per pass (the sample index lives
 * on the device). A view's samples are exactly those of rendering it alone with the
 * same seed (RNG keyed by seed, sample and pixel within the view). Unguided only.  */
typedef struct pg_batch pg_batch;
typedef struct pg_batch_config {
    uint32_t queue_paths;         /* paths per chunk (0: all views in one chunk)      */
    uint32_t trace_blocks_per_cu; /* traversal grid (0: as many as the LDS allows)    */
    uint32_t shade_blocks_per_cu; /* shade grid, blocks of 256 lanes (0: 3)       */
    uint32_t stack;               /* traversal stack entries (0: deepest BVH, rounded
                                     up to 16; must hold the deepest BVH)             */
    uint32_t use_graph;           /* replay a captured hipGraph per SPP               */
    uint32_t dyn_fetch;           /* persistent lanes with ray claims (see render params) */
    uint32_t samples_per_pass;    /* samples per pixel traced together in one pass (0: 1;
                                     at most 64, and with more than one at most 2^26
                                     paths per chunk); a render's samples beyond whole
                                     passes run as one shorter pass, and a render of
                                     fewer samples than a pass as one pass of its own
                                     size (replayed as a graph like the others); views
                                     are identical for every value                    */
    uint32_t stack_lds;           /* traversal stack entries per lane in LDS (see the
                                     render params; 0: 16)                            */
} pg_batch_config;
typedef struct pg_batch_stats {
    uint64_t paths, closest_rays, shadow_rays;
    double gpu_ms;                /* GPU time of the rendered passes                  */
    uint32_t views, scenes, chunks, graph_nodes;
} pg_batch_stats;
pg_status pg_batch_create(pg_ctx* ctx, pg_batch** out);
pg_status pg_batch_destroy(pg_batch* batch);
/* A scene as pg_scene_upload takes it (its BVH is built on the host here).        */
pg_status pg_batch_add_scene(pg_batch* batch, const float* positions, const float* normals,
                             const uint32_t* material_index, uint32_t n_tris,
                             const pg_material* materials, uint32_t n_materials,
                             const float* env_rgb, int32_t env_width, int32_t env_height,
                             uint32_t* scene_id);
/* Albedo textures of a batch scene: the arguments of pg_scene_textures. */
pg_status pg_batch_scene_textures(pg_batch* batch, uint32_t scene_id, const float* uv,
                                  const int32_t* material_texture, uint32_t n_materials, const float* texels,
                                  const uint64_t* texel_offset, const int32_t* tex_width,
                                  const int32_t* tex_height, uint32_t n_textures);
pg_status pg_batch_add_view(pg_batch* batch, uint32_t scene_id, const pg_camera_desc* camera,
                            uint32_t* view_id);
pg_status pg_batch_configure(pg_batch* batch, const pg_batch_config* cfg);
/* Times candidate configurations on one or two passes each and keeps the fastest,
 * then clears the films. It starts from chunks of about 2^20 paths (each chunk's
 * scenes then stay resident in L2 and the Infinity Cache while they are traced) with
 * the most samples per pass the 2^26-path state allows, and descends coordinate by
 * coordinate: chunk size with samples per pass, fewer samples per pass, graph, ray
 * claims, traversal and shade blocks per CU, LDS stack entries. Graph replay is kept
 * unless eager launches are more than 2% faster. */
pg_status pg_batch_autotune(pg_batch* batch, const pg_render_params* params, pg_batch_config* best,
                            double* ms_per_spp);
/* Learned guide of one scene of the batch (no reference counterpart; configuration 5
 * renders its views "guided by config 4's network with per-scene weights"): the
 * configuration the network was trained with and its FP32 parameters and hash grid
 * as pg_mlp_get_params returns them (17648 and 16 * 2^log2_table * 2 floats).
 * pg_batch_render with params->integrator = PG_INTEGRATOR_GUIDED_MLP then evaluates
 * each scene's network for its own paths (one-sample MIS with the BSDF, as a plain
 * PG_INTEGRATOR_GUIDED_MLP pass of pg_render_spp); scenes without a guide render
 * unguided. No training happens inside a batch.                                  */
pg_status pg_batch_set_guide(pg_batch* batch, uint32_t scene_id, const pg_mlp_config* cfg,
                             const float* params, const float* grid);
pg_status pg_batch_clear(pg_batch* batch);
/* params->integrator: PG_INTEGRATOR_PATH or PG_INTEGRATOR_GUIDED_MLP.             */
pg_status pg_batch_render(pg_batch* batch, const pg_render_params* params, uint32_t n_spp);
pg_status pg_batch_film_read(pg_batch* batch, uint32_t view_id, float* rgb_sum, uint32_t* spp);
pg_status pg_batch_stats_read(pg_batch* batch, pg_batch_stats* stats);

/* ------------------------------------------------------------------ host-side helpers */
/* Image metrics (Metrics.h:8-69): MSE, RMSE and RRMSE over width*height*3 values with
 * the reference's FP32 sequential sums (RRMSE divides by sqrt of the summed squared
 * truth + 1e-2, as Metrics.h:64-68 does), plus relMSE = mean (I-R)^2 / (R^2 + 1e-2)
 * in FP64. Mismatched sizes give -1 values and PG_ERR_INVALID (Metrics.h:10-13).   */
typedef struct { float mse, rmse, rrmse; double relmse; } pg_image_metrics;
pg_status pg_metrics(const float* test, const float* truth, uint64_t n_test, uint64_t n_truth,
                     int32_t width, int32_t height, pg_image_metrics* out);

/* Procedural scenes of the benchmark configurations (no reference counterpart:
 * the reference loads scene.json + .gem files, SceneLoader.h:222-269).
 * config 1: Cornell box; 2: open interior (~1.09M tris, env map);
 * 3: interior closed except a 2x2 window; 5: config-2 family for batched views. */
typedef struct pg_host_scene pg_host_scene;
pg_status pg_scene_generate(int32_t config, uint32_t seed, pg_host_scene** out);
pg_status pg_scene_sizes(const pg_host_scene* s, uint32_t* n_tris, uint32_t* n_materials,
                         int32_t* env_width, int32_t* env_height);
pg_status pg_scene_copy(const pg_host_scene* s, float* positions, float* normals,
                        uint32_t* material_index, pg_material* materials, float* env_rgb,
                        pg_camera_desc* camera);
pg_status pg_scene_free(pg_host_scene* s);

/* loadScene (SceneLoader.h:185-269): scene.json + .gem static meshes + PNG / Radiance
 * HDR textures of a reference scene directory, into a host scene. Materials follow
 * loadInstance (SceneLoader.h:104-181), one per instance; an unknown bsdf skips the
 * instance as the reference does; every material gets its reflectance texture (a
 * file that does not decode becomes the 1x1 white default, Imaging.h:36-42).
 * flags: PG_LOAD_SKIP_MISSING skips instances whose mesh file is missing or not a
 * GEM file (the reference exits instead). Errors: PG_ERR_INVALID with the reason in
 * pg_scene_load_error(). Images: 8-bit non-interlaced PNG, baseline and
 * progressive Huffman JPEG (8-bit, gray or YCbCr), Radiance HDR.                  */
enum { PG_LOAD_SKIP_MISSING = 1 };
pg_status pg_scene_load(const char* scene_dir, uint32_t flags, pg_host_scene** out);
const char* pg_scene_load_error(void);
pg_status pg_scene_texture_sizes(const pg_host_scene* s, uint32_t* n_textures, uint64_t* n_texels,
                                 uint32_t* n_warnings);
/* uv: n_tris*6 (zeros for procedural scenes); material_texture: n_materials (-1 for
 * procedural scenes); texels: n_texels*3; per texture: first texel, width, height. */
pg_status pg_scene_copy_textures(const pg_host_scene* s, float* uv, int32_t* material_texture, float* texels,
                                 uint64_t* texel_offset, int32_t* tex_width, int32_t* tex_height);
const char* pg_scene_warning(const pg_host_scene* s, uint32_t i);
/* Decode one image as Texture::load's decoder does (8-bit PNG / JPEG samples as
 * 0..255 floats, Radiance HDR as floats), `channels` values per pixel; rgb nullable. */
pg_status pg_image_read(const char* path, int32_t* width, int32_t* height, int32_t* channels, float* rgb);
/* RayTracer::savePNG (Renderer.h:751): 8-bit RGB PNG. */
pg_status pg_image_write_png(const char* path, const uint8_t* rgb, int32_t width, int32_t height);
/* RayTracer::saveHDR / Film::save (Renderer.h:750, Imaging.h:263-272): Radiance RGBE
 * of an RGB float image (pass the film divided by its SPP).                       */
pg_status pg_image_write_hdr(const char* path, const float* rgb, int32_t width, int32_t height);

/* Host BVH build alone (Geometry.h:419-446) — for topology parity checks. */
pg_status pg_bvh_build_host(const float* positions, uint32_t n_tris, uint32_t* n_nodes,
                            float* nodes_out /* nullable: 8*(2n-1) floats */,
                            int32_t* tri_index_out /* nullable: n */);

#ifdef __cplusplus
}
#endif
#endif /* PG_TRACER_H */
