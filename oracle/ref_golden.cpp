/*
 * ref_golden.cpp — C entry points over the reference's own hot-path classes.
 *
 * TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) from the
 * reference headers where they lie under /root/reference, into oracle/_ref/.
 * It is used in this container to write golden vectors (tests/golden/) that pin
 * the oracle restatement; it is never part of the product library.
 *
 * Build notes:
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
#include <sstream>
#include <fstream>
#include <iostream>
#include "ThirdParty/GamesEngineering/GEMLoader.h"  // self-contained; before the access widening
#define private public
#include "Geometry.h"
#include "Scene.h"
#include "Metrics.h"
#undef private
#include <map>
#include <string>
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
    bool layered = type >= 8;
    if (layered) type -= 8;
    switch (type) {
    case 0: b = new DiffuseBSDF(a); break;
    case 1: b = new MirrorBSDF(a); break;
    case 2: b = new ConductorBSDF(a, Colour(f[7], f[8], f[9]), Colour(f[10], f[11], f[12]), f[15]); break;
    case 3: b = new GlassBSDF(a, f[13], f[14]); break;
    case 5: b = new OrenNayarBSDF(a, f[15]); break;
    case 6: b = new PlasticBSDF(a, f[13], f[14], f[15]); break;
    default: b = new DielectricBSDF(a, f[13], f[14], f[15]); break;
    }
    b->emission = Colour(f[4], f[5], f[6]);
    if (layered) b = new LayeredBSDF(b, Colour(0.5f, 0.5f, 0.5f), 1.f, 1.3f, 1.f);
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
    if ((int)mat[0] % 8 != 4) b->invert(s, wi, iu, iv);  // DielectricBSDF::invert is empty
    out[11] = iu; out[12] = iv; out[13] = b->isPureSpecular() ? 1.f : 0.f;
    out[14] = b->isTwoSided() ? 1.f : 0.f;
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

/* Texture::sample (Imaging.h:82-108) on an RGB float texture. */
void ref_texture_sample(const float* rgb, int w, int h, const float* tuv, int n, float* out) {
    Texture t;
    t.width = w; t.height = h; t.channels = 3; t.alpha = nullptr;
    t.texels = new Colour[(size_t)w * h];
    for (size_t i = 0; i < (size_t)w * h; ++i) t.texels[i] = Colour(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    for (int i = 0; i < n; ++i) {
        Colour c = t.sample(tuv[2 * i], tuv[2 * i + 1]);
        out[3 * i] = c.r; out[3 * i + 1] = c.g; out[3 * i + 2] = c.b;
    }
}

/* Texture::load (Imaging.h:47-80) of one image file (stb_image): w, h, channels and
 * the RGB texels (w*h*3 floats, caller-sized via a first call with out == NULL). */
int ref_texture_load(const char* path, int* w, int* h, float* out) {
    Texture t;
    t.load(path);
    *w = t.width; *h = t.height;
    if (out)
        for (int i = 0; i < t.width * t.height; ++i) {
            out[3 * i] = t.texels[i].r; out[3 * i + 1] = t.texels[i].g; out[3 * i + 2] = t.texels[i].b;
        }
    return t.channels;
}

/* The scene-file path of SceneLoader.h:104-269 (loadInstance / loadScene) written
 * over the reference's own GEMScene / GEMModelLoader / Matrix / Triangle classes
 * (SceneLoader.h itself includes Renderer.h, which needs a Windows-only header).
 * Outputs per kept triangle: positions (9), normals (9), uv (6), material index;
 * per material: the 16-float record of the oracle (type + 8 when layered) and the
 * texture file name index; camera: from3 to3 up3 fov w h.                         */
struct RefLoaded {
    std::vector<float> pos, nrm, uv, mats, cam;
    std::vector<uint32_t> mi;
    std::vector<int32_t> mat_tex;
    std::vector<std::string> tex_names;
};

void* ref_load_scene(const char* dir) {
    std::string sceneName(dir);
    RefLoaded* r = new RefLoaded();
    GEMLoader::GEMScene gemscene;
    gemscene.load(sceneName + "/scene.json");
    int width = gemscene.findProperty("width").getValue(1920);
    int height = gemscene.findProperty("height").getValue(1080);
    float fov = gemscene.findProperty("fov").getValue(45.0f);
    Vec4 from, to, up;
    gemscene.findProperty("from").getValuesAsVector3(from.x, from.y, from.z);
    gemscene.findProperty("to").getValuesAsVector3(to.x, to.y, to.z);
    gemscene.findProperty("up").getValuesAsVector3(up.x, up.y, up.z);
    r->cam = {from.x, from.y, from.z, to.x, to.y, to.z, up.x, up.y, up.z, fov, (float)width, (float)height};
    std::map<std::string, int> texIndex;
    auto tex = [&](const std::string& f) {
        auto it = texIndex.find(f);
        if (it != texIndex.end()) return it->second;
        int k = (int)r->tex_names.size();
        texIndex[f] = k;
        r->tex_names.push_back(f);
        return k;
    };
    for (size_t ii = 0; ii < gemscene.instances.size(); ++ii) {
        GEMLoader::GEMInstance& instance = gemscene.instances[ii];
        GEMLoader::GEMModelLoader loader;
        std::vector<GEMLoader::GEMMesh> meshes;
        loader.load(sceneName + "/" + instance.meshFilename, meshes);
        std::string bsdf = instance.material.find("bsdf").getValue("");
        float m[16] = {0};
        bool ok = true;
        std::string refl = sceneName + "/" + instance.material.find("reflectance").getValue("");
        if (bsdf == "diffuse") { m[0] = 0; }
        else if (bsdf == "orennayar") { m[0] = 5; m[15] = instance.material.find("alpha").getValue(1.0f); }
        else if (bsdf == "glass") { m[0] = 3; m[13] = instance.material.find("intIOR").getValue(1.33f);
                                    m[14] = instance.material.find("extIOR").getValue(1.0f); }
        else if (bsdf == "mirror") { m[0] = 1; }
        else if (bsdf == "plastic") { m[0] = 6; m[13] = instance.material.find("intIOR").getValue(1.33f);
                                      m[14] = instance.material.find("extIOR").getValue(1.0f);
                                      m[15] = instance.material.find("roughness").getValue(1.0f); }
        else if (bsdf == "dielectric") { m[13] = instance.material.find("intIOR").getValue(1.33f);
                                         m[14] = instance.material.find("extIOR").getValue(1.0f);
                                         m[15] = instance.material.find("roughness").getValue(1.0f);
                                         m[0] = m[15] < 0.001f ? 3 : 4; if (m[0] == 3) m[15] = 0.f; }
        else if (bsdf == "conductor") { m[0] = 2;
            instance.material.find("eta").getValuesAsVector3(m[7], m[8], m[9]);
            instance.material.find("k").getValuesAsVector3(m[10], m[11], m[12]);
            m[15] = instance.material.find("roughness").getValue(1.0f); }
        else ok = false;
        if (!ok) continue;
        if (instance.material.find("emission").getValue("") != "")
            instance.material.find("emission").getValuesAsVector3(m[4], m[5], m[6]);
        if (instance.material.find("coatingThickness").getValue(0) > 0) m[0] += 8;
        for (int k = 0; k < 16; ++k) r->mats.push_back(m[k]);
        r->mat_tex.push_back(tex(refl));
        uint32_t materialIndex = (uint32_t)(r->mat_tex.size() - 1);
        std::vector<Vertex> vertices;
        std::vector<unsigned int> indices;
        Matrix transform;
        memcpy(transform.m, instance.w.m, 16 * sizeof(float));
        Matrix vecTransform = transform.invert();
        vecTransform = vecTransform.transpose();
        for (size_t i = 0; i < meshes.size(); i++) {
            for (size_t n = 0; n < meshes[i].verticesStatic.size(); n++) {
                Vertex v;
                v.p.x = meshes[i].verticesStatic[n].position.x;
                v.p.y = meshes[i].verticesStatic[n].position.y;
                v.p.z = meshes[i].verticesStatic[n].position.z;
                v.normal.x = meshes[i].verticesStatic[n].normal.x;
                v.normal.y = meshes[i].verticesStatic[n].normal.y;
                v.normal.z = meshes[i].verticesStatic[n].normal.z;
                v.p = transform.mulPoint(v.p);
                v.normal = vecTransform.mulVec(v.normal);
                v.normal = v.normal.normalize();
                v.u = meshes[i].verticesStatic[n].u;
                v.v = meshes[i].verticesStatic[n].v;
                vertices.push_back(v);
            }
            int offset = (int)indices.size();
            for (size_t n = 0; n < meshes[i].indices.size(); n++) indices.push_back(offset + meshes[i].indices[n]);
        }
        for (size_t i = 0; i + 2 < indices.size(); i += 3) {
            Triangle t;
            t.init(vertices[indices[i]], vertices[indices[i + 1]], vertices[indices[i + 2]], materialIndex);
            if (t.area > 0) {
                for (int k = 0; k < 3; ++k) {
                    r->pos.push_back(t.vertices[k].p.x); r->pos.push_back(t.vertices[k].p.y); r->pos.push_back(t.vertices[k].p.z);
                    r->nrm.push_back(t.vertices[k].normal.x); r->nrm.push_back(t.vertices[k].normal.y);
                    r->nrm.push_back(t.vertices[k].normal.z);
                    r->uv.push_back(t.vertices[k].u); r->uv.push_back(t.vertices[k].v);
                }
                r->mi.push_back(materialIndex);
            }
        }
    }
    return r;
}
int ref_loaded_sizes(void* p, int* n_mats, int* n_tex) {
    RefLoaded* r = (RefLoaded*)p;
    *n_mats = (int)r->mat_tex.size(); *n_tex = (int)r->tex_names.size();
    return (int)r->mi.size();
}
void ref_loaded_copy(void* p, float* pos, float* nrm, float* uv, uint32_t* mi, float* mats, int32_t* mat_tex,
                     float* cam) {
    RefLoaded* r = (RefLoaded*)p;
    memcpy(pos, r->pos.data(), r->pos.size() * 4); memcpy(nrm, r->nrm.data(), r->nrm.size() * 4);
    memcpy(uv, r->uv.data(), r->uv.size() * 4); memcpy(mi, r->mi.data(), r->mi.size() * 4);
    memcpy(mats, r->mats.data(), r->mats.size() * 4); memcpy(mat_tex, r->mat_tex.data(), r->mat_tex.size() * 4);
    memcpy(cam, r->cam.data(), r->cam.size() * 4);
}
const char* ref_loaded_texture_name(void* p, int i) { return ((RefLoaded*)p)->tex_names[i].c_str(); }
void ref_loaded_free(void* p) { delete (RefLoaded*)p; }

/* Film::tonemap (Imaging.h:220-244) of one accumulated pixel. */
void ref_tonemap(const float* rgb_sum, int n, int spp, float exposure, float lmax, uint8_t* out) {
    Film f;
    f.width = (unsigned)n; f.height = 1; f.SPP = spp;
    f.film = new Colour[n];
    for (int i = 0; i < n; ++i) f.film[i] = Colour(rgb_sum[3 * i], rgb_sum[3 * i + 1], rgb_sum[3 * i + 2]);
    for (int i = 0; i < n; ++i) {
        unsigned char r, g, b;
        f.tonemap(i, 0, r, g, b, exposure, lmax);
        out[3 * i] = r; out[3 * i + 1] = g; out[3 * i + 2] = b;
    }
    delete[] f.film;
}

/* A whole Scene over the reference's classes (Scene.h:84-137): triangles with
 * per-vertex normals (uv 0), the materials as BSDFs (16-float records), the
 * background (EnvironmentMap over an RGB texture, or a black BackgroundColour),
 * Scene::init + Scene::build (BVH and the light list).                          */
struct RefFull {
    Scene scene;
};

void* ref_full_create(const float* pos, const float* nrm, const uint32_t* mi, uint32_t n, const float* mats,
                      uint32_t n_mats, const float* env_rgb, int env_w, int env_h) {
    RefFull* r = new RefFull();
    std::vector<Triangle> tris(n);
    for (uint32_t i = 0; i < n; ++i) {
        Vertex v[3];
        for (int k = 0; k < 3; ++k) {
            v[k].p = Vec4(pos[9 * i + 3 * k], pos[9 * i + 3 * k + 1], pos[9 * i + 3 * k + 2]);
            v[k].normal = Vec4(nrm[9 * i + 3 * k], nrm[9 * i + 3 * k + 1], nrm[9 * i + 3 * k + 2]);
            v[k].u = 0.f; v[k].v = 0.f;
        }
        tris[i].init(v[0], v[1], v[2], mi[i]);
    }
    std::vector<BSDF*> bs;
    for (uint32_t i = 0; i < n_mats; ++i) bs.push_back(makeBSDF(mats + 16 * (size_t)i));
    Light* bg = nullptr;
    if (env_rgb && env_w > 0 && env_h > 0) bg = (Light*)ref_env_create(env_rgb, env_w, env_h);
    else bg = new BackgroundColour(Colour(0.f, 0.f, 0.f));
    r->scene.init(tris, bs, bg);
    r->scene.build();
    return r;
}
void ref_full_destroy(void* p) { delete (RefFull*)p; }

/* The light list: kind (0 area, 1 environment) and the area light's triangle. */
int ref_full_lights(void* p, int32_t* kind, int32_t* tri) {
    Scene& s = ((RefFull*)p)->scene;
    for (size_t i = 0; i < s.lights.size(); ++i) {
        bool area = s.lights[i]->isArea();
        if (kind) kind[i] = area ? 0 : 1;
        if (tri) tri[i] = area ? (int32_t)(((AreaLight*)s.lights[i])->triangle - s.triangles.data()) : -1;
    }
    return (int)s.lights.size();
}

/* AreaLight::sample / evaluate / PDF / normal (Lights.h:28-73; Triangle::sample,
 * Geometry.h:139-150) of area light `light` with random numbers u2 and the
 * evaluation direction wi. out: point[3], pdf, emission[3], evaluate(wi)[3],
 * PDF(wi), normal[3] (14 floats).                                                */
void ref_area_light(void* p, int light, const float* u2, const float* wi, float* out) {
    Scene& s = ((RefFull*)p)->scene;
    Light* L = s.lights[light];
    ListSampler ls; ls.u[0] = u2[0]; ls.u[1] = u2[1]; ls.n = 2;
    ShadingData sd;
    Colour e; float pdf = 0.f;
    Vec4 x = L->sample(sd, &ls, e, pdf);
    Vec4 w(wi[0], wi[1], wi[2]);
    Colour ev = L->evaluate(w);
    Vec4 nn = L->normal(sd, w);
    float o[14] = {x.x, x.y, x.z, pdf, e.r, e.g, e.b, ev.r, ev.g, ev.b, L->PDF(sd, w), nn.x, nn.y, nn.z};
    std::memcpy(out, o, sizeof(o));
}

/* Scene::sampleLight (Scene.h:108-119) for n random numbers: index and pmf. */
void ref_sample_light(void* p, const float* u, uint32_t n, int32_t* idx, float* pmf) {
    Scene& s = ((RefFull*)p)->scene;
    for (uint32_t i = 0; i < n; ++i) {
        ListSampler ls; ls.u[0] = u[i]; ls.n = 1;
        float pm = 0.f;
        Light* L = s.sampleLight(&ls, pm);
        int k = -1;
        for (size_t j = 0; j < s.lights.size(); ++j)
            if (s.lights[j] == L) k = (int)j;
        idx[i] = k; pmf[i] = pm;
    }
}

/* Scene::traverse + Scene::calculateShadingData (Scene.h:152-176) for n rays.
 * out per ray (24 floats): x[3], gNormal[3], sNormal[3], wo[3], tu, tv, t,
 * frame.u[3], frame.v[3], frame.w[3] (zeros where the ray misses).               */
void ref_shading_data(void* p, const float* rays, uint32_t n, float* out) {
    Scene& s = ((RefFull*)p)->scene;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 6 * (size_t)i;
        Ray ray(Vec4(r[0], r[1], r[2]), Vec4(r[3], r[4], r[5]));
        IntersectionData h = s.traverse(ray);
        ShadingData sd = s.calculateShadingData(h, ray);
        float* o = out + 24 * (size_t)i;
        std::memset(o, 0, 24 * sizeof(float));
        o[9] = sd.wo.x; o[10] = sd.wo.y; o[11] = sd.wo.z; o[14] = sd.t;
        if (h.t < FLT_MAX) {
            o[0] = sd.x.x; o[1] = sd.x.y; o[2] = sd.x.z;
            o[3] = sd.gNormal.x; o[4] = sd.gNormal.y; o[5] = sd.gNormal.z;
            o[6] = sd.sNormal.x; o[7] = sd.sNormal.y; o[8] = sd.sNormal.z;
            o[12] = sd.tu; o[13] = sd.tv;
            o[15] = sd.frame.u.x; o[16] = sd.frame.u.y; o[17] = sd.frame.u.z;
            o[18] = sd.frame.v.x; o[19] = sd.frame.v.y; o[20] = sd.frame.v.z;
            o[21] = sd.frame.w.x; o[22] = sd.frame.w.y; o[23] = sd.frame.w.z;
        }
    }
}

/* Scene::visible (Scene.h:139-146) on the whole scene. */
void ref_full_visible(void* p, const float* p1, const float* p2, uint32_t n, uint8_t* vis) {
    Scene& s = ((RefFull*)p)->scene;
    for (uint32_t i = 0; i < n; ++i)
        vis[i] = s.visible(Vec4(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]), Vec4(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]))
                     ? 1 : 0;
}

/* Film::splat (Imaging.h:186-217) with the BoxFilter (Imaging.h:147-151) of n
 * samples (x, y, L), in order, onto a zeroed w x h film; out: the film sums.     */
void ref_film_splat(int w, int h, const float* xy, const float* L, uint32_t n, float* out) {
    Film f;
    BoxFilter box;
    f.init(w, h, &box);
    for (uint32_t i = 0; i < n; ++i) f.splat(xy[2 * i], xy[2 * i + 1], Colour(L[3 * i], L[3 * i + 1], L[3 * i + 2]));
    for (int i = 0; i < w * h; ++i) { out[3 * i] = f.film[i].r; out[3 * i + 1] = f.film[i].g; out[3 * i + 2] = f.film[i].b; }
    delete[] f.film;
}

/* stbi_write_hdr of the mean image, as Film::save (Imaging.h:263-272). */
int ref_write_hdr(const char* path, const float* rgb, int w, int h) { return stbi_write_hdr(path, w, h, 3, rgb); }

}  // extern "C"
