/*
 * pg_oracle.h — C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a clean-room CPU restatement of the
 * reference renderer (IsaGeriler/PathGuidingDissertation) used as the checker
 * for the HIP product path and as the CPU baseline in bench.py. Nothing in the
 * product library links, loads or calls it.
 *
 * Arithmetic is FP32 with FMA contraction off (-ffp-contract=off), following the
 * reference's operation order; the film can accumulate in FP64 for oracle images.
 */
#ifndef PG_ORACLE_H
#define PG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Material record: 16 floats
 *   [0] type (0 diffuse, 1 mirror, 2 conductor, 3 glass, 4 dielectric,
 *       5 Oren-Nayar, 6 plastic; + 8 = LayeredBSDF over that base)
 *   [1..3] albedo  [4..6] emission  [7..9] eta  [10..12] k
 *   [13] intIOR  [14] extIOR  [15] roughness (Oren-Nayar: sigma)           */
#define ORC_MAT_FLOATS 16

/* cam: from[3], to[3], up[3], fov_deg, width, height (12 floats)
 * env_rgb: env_w*env_h*3 floats (lat-long, y-up) or NULL for a black background */
orc_scene* orc_scene_create(const float* pos, const float* nrm, const uint32_t* mat_idx,
                            uint32_t n_tris, const float* mats, uint32_t n_mats,
                            const float* env_rgb, int32_t env_w, int32_t env_h,
                            const float* cam);
void orc_scene_destroy(orc_scene* s);
/* Albedo textures (Imaging.h:22-145) and per-vertex uv (Vertex::u, v): uv n_tris*6
 * (u0 v0 u1 v1 u2 v2), mat_tex per material (-1: constant albedo), texels RGB of
 * every texture concatenated (texture i starts at texel offset[i], w[i] x h[i]). */
void orc_scene_set_textures(orc_scene* s, const float* uv, const int32_t* mat_tex, uint32_t n_mats,
                            const float* texels, const uint64_t* offset, const int32_t* w, const int32_t* h,
                            uint32_t n_tex);
/* Texture::sample (Imaging.h:82-108) at n (tu, tv) pairs of one w x h RGB texture. */
void orc_texture_sample(const float* texels, int32_t w, int32_t h, const float* tuv, uint32_t n, float* rgb);

/* BVH (binned SAH, Geometry.h:293-446). nodes: 8 floats each
 * (min.xyz, leftFirst (as int bits), max.xyz, used (as int bits)). */
uint32_t orc_bvh_node_count(const orc_scene* s);
void orc_bvh_export(const orc_scene* s, float* nodes, int32_t* tri_idx);
int32_t orc_bvh_max_depth(const orc_scene* s);

/* Camera matrices: inv_proj[16], camera[16], origin[3] */
void orc_camera_export(const orc_scene* s, float* inv_proj, float* cam, float* origin);
void orc_generate_rays(const orc_scene* s, const float* pxy, uint32_t n, float* rays6);

/* Closest hit (Geometry.h:448-512). rays: n*6 (o.xyz, d.xyz). Counters nullable. */
void orc_trace_closest(const orc_scene* s, const float* rays, uint32_t n, uint32_t* id,
                       float* t, float* alpha, float* beta, float* gamma,
                       uint32_t* n_nodes, uint32_t* n_tris, int32_t n_threads);
/* Any hit (Geometry.h:514-567). vis[i] = 1 if unoccluded within (1e-4, maxT). */
void orc_trace_shadow(const orc_scene* s, const float* rays, const float* max_t, uint32_t n,
                      uint8_t* vis, uint32_t* n_nodes, uint32_t* n_tris, int32_t n_threads);
/* Brute-force closest hit over all triangles (lowest index wins exact ties). */
void orc_trace_brute(const orc_scene* s, const float* rays, uint32_t n, uint32_t* id, float* t);

/* Render parameters */
typedef struct {
    uint32_t spp;          /* samples per pixel rendered by this call          */
    uint32_t spp_offset;   /* index of the first sample (RNG stream offset)    */
    uint32_t seed;
    uint32_t max_depth;    /* Renderer.h:568 uses 8                            */
    uint32_t rr_depth;     /* Russian roulette when depth > rr_depth (3)       */
    int32_t integrator;    /* 0 path tracer (NEE+MIS), 1 guided record pass,
                              2 record pass whose Li carries NEE (pixelColour) */
    int32_t sampler;       /* 0 counter hash, 1 std::mt19937 per thread        */
    int32_t accum_f64;     /* 1: film sum in FP64                               */
    int32_t n_threads;     /* 0: hardware_concurrency                           */
    int32_t record;        /* integrator 1: collect PathVertex records          */
    int32_t env_mode;      /* 0: Distribution2D (Sampling.h:166-209, 2 draws);
                              1: alias table over the same texel pmf (4 draws),
                                 the sampling the device path uses             */
    int32_t env_visibility; /* 0: shadow ray to sceneCentre + wi * sceneRadius
                                  (Renderer.h:331-332); 1: along wi out of the
                                  bounding sphere                               */
    int32_t crop[4];       /* x0, y0, width, height of the pixels traced (the
                              rest of the film stays untouched); width 0: all  */
} orc_render_params;

typedef struct {
    uint64_t closest_rays, shadow_rays, paths;
    double seconds;
    uint64_t n_records;
} orc_render_stats;

/* film: W*H*3 doubles, accumulated (sum over samples, not divided by SPP).  */
int orc_render(orc_scene* s, const orc_render_params* p, double* film, orc_render_stats* st);
/* Records gathered by the last record-pass render: 15 floats each
 * (pos.xyz, normal.xyz, wi.xyz, Li.rgb, pixel, sample, depth).              */
uint64_t orc_records_count(const orc_scene* s);
void orc_records_export(const orc_scene* s, float* out);

/* Counter-hash RNG: value for (seed, sample, pixel, dim). */
float orc_rng(uint32_t seed, uint32_t sample, uint32_t pixel, uint32_t dim);

/* BSDF unit access. sd: x[3], wo[3], sN[3], gN[3] (frame built from sN).
 * u: random numbers consumed in order (up to 3).                           */
void orc_bsdf_sample(const float* mat, const float* sd, const float* u, float* wi,
                     float* f_rgb, float* pdf);
void orc_bsdf_eval(const float* mat, const float* sd, const float* wi, float* f_rgb, float* pdf);
void orc_bsdf_invert(const float* mat, const float* sd, const float* wi, float* uv);
/* BSDF::isPureSpecular and isTwoSided of a material record. */
void orc_bsdf_flags(const float* mat, int32_t* specular, int32_t* two_sided);
/* Film::tonemap (Imaging.h:220-244) of n accumulated pixels to 8-bit RGB. */
void orc_tonemap(const float* rgb_sum, uint64_t n, int32_t spp, float exposure, float lmax_white, uint8_t* out);
float orc_fresnel_dielectric(float cos_theta, float int_ior, float ext_ior);
void orc_fresnel_conductor(float cos_theta, const float* eta, const float* k, float* out);
float orc_ggx_d(const float* h, float alpha);
float orc_ggx_g(const float* wi, const float* wo, float alpha);

/* Environment map (Lights.h:128-228, Sampling.h:166-209) */
void orc_env_sample(const orc_scene* s, float ru, float rv, float* wi, float* rgb, float* pdf);
/* Alias-table sample from four numbers (texel, acceptance, in-texel u, v). */
void orc_env_sample_alias(const orc_scene* s, const float* u4, float* wi, float* rgb, float* pdf);
/* Alias table export: prob, alias, pdf (w*h each). */
void orc_env_tables(const orc_scene* s, float* prob, uint32_t* alias, float* pdf);
/* Metrics::calculateMSE / RMSE / RRMSE (Metrics.h:8-69): FP32, sequential sums over
 * width*height*3 values; size mismatch gives -1. out: mse, rmse, rrmse.           */
void orc_metrics(const float* test, const float* truth, uint64_t n_test, uint64_t n_truth, int width, int height,
                 float* out);
void orc_env_eval(const orc_scene* s, const float* wi, float* rgb, float* pdf);
float orc_env_marginal_integral(const orc_scene* s);
void orc_scene_bounds(const orc_scene* s, float* centre, float* radius);

/* NEE constituents, one call per unit (parity aids):
 * the light list (Scene.h:93-100, 133-136): kind 0 area (tri = its triangle), 1 env;
 * returns the count. */
int32_t orc_lights(const orc_scene* s, int32_t* kind, int32_t* tri);
/* AreaLight::sample / evaluate / PDF / normal (Lights.h:28-73, Triangle::sample
 * Geometry.h:139-150) of area light `light` with random numbers u2 and direction wi.
 * out (14): point[3], pdf, emission[3], evaluate(wi)[3], PDF, normal[3].          */
void orc_area_light(const orc_scene* s, int32_t light, const float* u2, const float* wi, float* out);
/* Scene::sampleLight (Scene.h:108-119) for n random numbers: light index, pmf. */
void orc_sample_light(const orc_scene* s, const float* u, uint32_t n, int32_t* idx, float* pmf);
/* Scene::calculateShadingData (Scene.h:152-176) of each ray's closest hit. out per ray
 * (24): x[3], gN[3], sN[3], wo[3], tu, tv, t, frame u[3], v[3], w[3] (misses: wo, t). */
void orc_shading_data(const orc_scene* s, const float* rays, uint32_t n, float* out);
/* Scene::visible (Scene.h:139-146). */
void orc_visible(const orc_scene* s, const float* p1, const float* p2, uint32_t n, uint8_t* vis);
/* Film::splat with the BoxFilter (Imaging.h:147-151, 186-217) of n samples (x, y, L)
 * in order, added into a w x h x 3 FP32 film. */
void orc_film_splat(int32_t w, int32_t h, const float* xy, const float* L, uint32_t n, float* film);

/* Point BVH over records (Renderer.h:71-234): longest-axis median split,
 * leaf <= 8. Returns node count; stats: nodes, leaves, min depth, max depth. */
int32_t orc_point_bvh_stats(const float* pos, uint32_t n, int32_t* stats);
/* Brute-force kNN over points (ties by lower index). idx: nq*k, d2: nq*k. */
void orc_knn_brute(const float* pts, uint32_t n, const float* q, uint32_t nq, uint32_t k,
                   uint32_t* idx, float* d2, int32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif
