/*
 * ref_golden.cpp — C entry points over the reference's own hot-path classes.
 *
 * TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) from the
 * reference headers where they lie under /root/reference, into oracle/_ref/.
 * It is used in this container to write golden vectors (tests/golden/) that pin
 * the oracle restatement; it is never part of the product library.
 *
 * Build notes (recorded in DESIGN.md):
 *   * libstdc++ 11 does not declare std::cosf/sinf/tanf/powf/sqrtf, which C++17
 *     <cmath> specifies and MSVC provides; the alias below names the same C
 *     functions, so no arithmetic changes.
 *   * `private` is widened to reach BVHNode's node array and triangle order
 *     (Geometry.h:269-272) for topology export; no reference code is altered.
 *   * Renderer.h / main.cpp are not built: they need the Windows-only
 *     ThirdParty/GamesEngineering/GamesEngineeringBase.h.
 */
#include <cmath>
#include <cfloat>
#include <climits>
#include <cstdio>
#include <memory>
#include <cstdint>
#include <cstring>
namespace std { using ::cosf; using ::sinf; using ::tanf; using ::powf; using ::sqrtf; }
#define private public
#include "Geometry.h"
#include "Scene.h"
#include "Metrics.h"
#undef private
#include <vector>

namespace {
struct ListSampler : Sampler {
    float u[4]; int n = 0, c = 0;
    float next() { return c < n ? u[c++] : 0.5f; }
};

Texture* constTexture(float r, float g, float b) {
    Texture* t = new Texture();
    t->width = 1; t->height = 1; t->channels = 3;
    t->texels = new Colour[1];
    t->texels[0] = Colour(r, g, b);
    t->alpha = nullptr;
    return t;
}

BSDF* makeBSDF(const float* f) {
    Texture* a = constTexture(f[1], f[2], f[3]);
    int type = (int)f[0];
    BSDF* b = nullptr;
    switch (type) {
    case 0: b = new DiffuseBSDF(a); break;
    case 1: b = new MirrorBSDF(a); break;
    case 2: b = new ConductorBSDF(a, Colour(f[7], f[8], f[9]), Colour(f[10], f[11], f[12]), f[15]); break;
    case 3: b = new GlassBSDF(a, f[13], f[14]); break;
    default: b = new DielectricBSDF(a, f[13], f[14], f[15]); break;
    }
    b->emission = Colour(f[4], f[5], f[6]);
    return b;
}

ShadingData makeSD(const float* sd, BSDF* b) {
    ShadingData s;
    s.x = Vec4(sd[0], sd[1], sd[2]);
    s.wo = Vec4(sd[3], sd[4], sd[5]);
    s.sNormal = Vec4(sd[6], sd[7], sd[8]);
    s.gNormal = Vec4(sd[9], sd[10], sd[11]);
    s.tu = 0.f; s.tv = 0.f; s.t = 1.f;
    s.frame.fromVector(s.sNormal);
    s.bsdf = b;
    return s;
}

struct RefScene {
    std::vector<Triangle> tris;
    BVHNode bvh;
};
}  // namespace

extern "C" {

void* ref_scene_create(const float* pos, const float* nrm, uint32_t n) {
    RefScene* s = new RefScene();
    s->tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        Vertex v[3];
        for (int k = 0; k < 3; ++k) {
            v[k].p = Vec4(pos[9 * i + 3 * k], pos[9 * i + 3 * k + 1], pos[9 * i + 3 * k + 2]);
            v[k].normal = Vec4(nrm[9 * i + 3 * k], nrm[9 * i + 3 * k + 1], nrm[9 * i + 3 * k + 2]);
            v[k].u = 0.f; v[k].v = 0.f;
        }
        s->tris[i].init(v[0], v[1], v[2], 0);
    }
    s->bvh.build(s->tris);
    return s;
}
void ref_scene_destroy(void* p) { delete (RefScene*)p; }
uint32_t ref_bvh_node_count(void* p) { return (uint32_t)((RefScene*)p)->bvh.bvhBuildNode.size(); }
void ref_bvh_export(void* p, float* nodes, int32_t* idx) {
    RefScene* s = (RefScene*)p;
    for (size_t i = 0; i < s->bvh.bvhBuildNode.size(); ++i) {
        const BVHBuildNode& n = s->bvh.bvhBuildNode[i];
        float* o = nodes + 8 * i;
        o[0] = n.bounds.min.x; o[1] = n.bounds.min.y; o[2] = n.bounds.min.z; std::memcpy(o + 3, &n.leftFirst, 4);
        o[4] = n.bounds.max.x; o[5] = n.bounds.max.y; o[6] = n.bounds.max.z; std::memcpy(o + 7, &n.used, 4);
    }
    for (size_t i = 0; i < s->bvh.triangleIndexes.size(); ++i) idx[i] = s->bvh.triangleIndexes[i];
}
void ref_trace_closest(void* p, const float* rays, uint32_t n, uint32_t* id, float* t, float* a,
                       float* b, float* g) {
    RefScene* s = (RefScene*)p;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 6 * (size_t)i;
        Ray ray(Vec4(r[0], r[1], r[2]), Vec4(r[3], r[4], r[5]));
        IntersectionData h = s->bvh.traverse(ray, s->tris);
        id[i] = h.ID; t[i] = h.t; a[i] = h.alpha; b[i] = h.beta; g[i] = h.gamma;
    }
}
void ref_trace_shadow(void* p, const float* rays, const float* maxT, uint32_t n, uint8_t* vis) {
    RefScene* s = (RefScene*)p;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 6 * (size_t)i;
        Ray ray(Vec4(r[0], r[1], r[2]), Vec4(r[3], r[4], r[5]));
        vis[i] = s->bvh.traverseVisible(ray, s->tris, maxT[i]) ? 1 : 0;
    }
}
/* Scene::visible (Scene.h:139-146) on the same BVH. */
void ref_visible(void* p, const float* p1, const float* p2, uint32_t n, uint8_t* vis) {
    RefScene* s = (RefScene*)p;
    for (uint32_t i = 0; i < n; ++i) {
        Vec4 a(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]), b(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]);
        Ray ray;
        Vec4 dir = b - a;
        float maxT = dir.length() - (2.f * EPSILON);
        dir = dir.normalize();
        ray.init(a + (dir * EPSILON), dir);
        vis[i] = s->bvh.traverseVisible(ray, s->tris, maxT) ? 1 : 0;
    }
}

/* Camera (SceneLoader.h:226-244, Scene.h:25-57). cam: from3,to3,up3,fov,w,h */
void ref_camera(const float* cam, const float* pxy, uint32_t n, float* inv_proj, float* camm,
                float* rays) {
    int W = (int)cam[10], H = (int)cam[11];
    Matrix P = Matrix::perspective(0.001f, 10000.0f, (float)W / (float)H, cam[9]);
    Vec4 from(cam[0], cam[1], cam[2]), to(cam[3], cam[4], cam[5]), up(cam[6], cam[7], cam[8]);
    Matrix V = Matrix::lookAt(from, to, up);
    V = V.invert();
    Camera c;
    c.init(P, W, H);
    c.updateView(V);
    std::memcpy(inv_proj, c.inverseProjectionMatrix.m, 64);
    std::memcpy(camm, c.camera.m, 64);
    for (uint32_t i = 0; i < n; ++i) {
        Ray r = c.generateRay(pxy[2 * i], pxy[2 * i + 1]);
        float* o = rays + 6 * (size_t)i;
        o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = r.dir.x; o[4] = r.dir.y; o[5] = r.dir.z;
    }
}

void ref_bsdf(const float* mat, const float* sd, const float* u, const float* wi_eval, float* out) {
    /* out: [0..2] sampled wi, [3..5] f, [6] pdf, [7..9] eval f at wi_eval, [10] pdf at wi_eval,
            [11..12] invert(sampled wi), [13] isPureSpecular                              */
    BSDF* b = makeBSDF(mat);
    ShadingData s = makeSD(sd, b);
    ListSampler ls; ls.u[0] = u[0]; ls.u[1] = u[1]; ls.u[2] = u[2]; ls.n = 3;
    Colour f; float pdf = 0.f;
    Vec4 wi = b->sample(s, &ls, f, pdf);
    out[0] = wi.x; out[1] = wi.y; out[2] = wi.z; out[3] = f.r; out[4] = f.g; out[5] = f.b; out[6] = pdf;
    Vec4 we(wi_eval[0], wi_eval[1], wi_eval[2]);
    Colour fe = b->evaluate(s, we);
    out[7] = fe.r; out[8] = fe.g; out[9] = fe.b; out[10] = b->PDF(s, we);
    float iu = 0.f, iv = 0.f;
    if ((int)mat[0] != 4) b->invert(s, wi, iu, iv);
    out[11] = iu; out[12] = iv; out[13] = b->isPureSpecular() ? 1.f : 0.f;
}

float ref_fresnel_dielectric(float c, float i, float e) { return ShadingHelper::fresnelDielectric(c, i, e); }
void ref_fresnel_conductor(float c, const float* eta, const float* k, float* out) {
    Colour r = ShadingHelper::fresnelConductor(c, Colour(eta[0], eta[1], eta[2]), Colour(k[0], k[1], k[2]));
    out[0] = r.r; out[1] = r.g; out[2] = r.b;
}
float ref_ggx_d(const float* h, float a) { return ShadingHelper::Dggx(Vec4(h[0], h[1], h[2]), a); }
float ref_ggx_g(const float* wi, const float* wo, float a) {
    return ShadingHelper::Gggx(Vec4(wi[0], wi[1], wi[2]), Vec4(wo[0], wo[1], wo[2]), a);
}

void* ref_env_create(const float* rgb, int w, int h) {
    Texture* t = new Texture();
    t->width = w; t->height = h; t->channels = 3; t->alpha = nullptr;
    t->texels = new Colour[(size_t)w * h];
    for (size_t i = 0; i < (size_t)w * h; ++i) t->texels[i] = Colour(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    return new EnvironmentMap(t);
}
void ref_env_destroy(void* p) { delete (EnvironmentMap*)p; }
void ref_env_sample(void* p, float ru, float rv, float* out) {
    /* out: wi[3], rgb[3], pdf */
    EnvironmentMap* e = (EnvironmentMap*)p;
    ListSampler ls; ls.u[0] = ru; ls.u[1] = rv; ls.n = 2;
    ShadingData sd; Colour c; float pdf = 0.f;
    Vec4 wi = e->sample(sd, &ls, c, pdf);
    out[0] = wi.x; out[1] = wi.y; out[2] = wi.z; out[3] = c.r; out[4] = c.g; out[5] = c.b; out[6] = pdf;
}
void ref_env_eval(void* p, const float* wi, float* out) {
    EnvironmentMap* e = (EnvironmentMap*)p;
    ShadingData sd;
    Vec4 w(wi[0], wi[1], wi[2]);
    Colour c = e->evaluate(w);
    out[0] = c.r; out[1] = c.g; out[2] = c.b; out[3] = e->PDF(sd, w);
}
float ref_env_power(void* p) { return ((EnvironmentMap*)p)->totalIntegratedPower(); }

void ref_metrics(const float* test, const float* truth, int n_test, int n_truth, int w, int h, float* out) {
    std::vector<float> a(test, test + n_test), b(truth, truth + n_truth);
    Metrics m;
    out[0] = m.calculateMSE(a, b, w, h);
    out[1] = m.calculateRMSE(a, b, w, h);
    out[2] = m.calculateRRMSE(a, b, w, h);
}

}  // extern "C"
