// host_prep.cpp — host preparation of the device scene.
//
// The BVH is built on the host in the reference's exact topology and triangle
// order (binned SAH, 16 bins, leaf <= 4, Geometry.h:293-446) so that the device
// traversal reproduces the reference's first-visited tie-breaking. Subtrees are
// built in parallel and renumbered afterwards into the reference's node order
// (children allocated as a pair when their parent splits, depth-first).
#include "host_prep.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

namespace pg {
namespace {

inline HV3 vmin3(const HV3& a, const HV3& b) { return HV3(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
inline HV3 vmax3(const HV3& a, const HV3& b) { return HV3(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }
inline HV3 cross3(const HV3& a, const HV3& b) {
    return HV3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
inline float dot3(const HV3& a, const HV3& b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
inline float len3(const HV3& a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
inline HV3 norm3(const HV3& a) { float l = 1.f / len3(a); return HV3(a.x * l, a.y * l, a.z * l); }

struct Box {
    HV3 mn{FLT_MAX, FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const HV3& p) { mx = vmax3(mx, p); mn = vmin3(mn, p); }
    // AABB::area (Geometry.h:212-215)
    float area() const {
        HV3 s = mx - mn;
        return (((s.x * s.y) + (s.y * s.z)) + (s.x * s.z)) * 2.f;
    }
};

struct TmpNode {
    Box b;
    int32_t first = 0, count = 0;
    int32_t child = -1;  // index of the left child in the temporary array (pair)
    uint32_t depth = 0;
};

struct Builder {
    const float* pos;
    uint32_t n;
    std::vector<HV3> cen, tmn, tmx;
    std::vector<int32_t> idx;
    std::vector<TmpNode> tmp;
    std::atomic<int32_t> next{1};
    std::atomic<uint32_t> max_depth{0};

    HV3 vert(uint32_t t, int k) const { const float* p = pos + 9 * (size_t)t + 3 * k; return HV3(p[0], p[1], p[2]); }

    // BVHNode::updateBounds (Geometry.h:274-287); extending by the triangle's own
    // min/max is identical to extending by its three vertices.
    Box range_bounds(int32_t first, int32_t count) const {
        Box b;
        for (int32_t k = 0; k < count; ++k) {
            int32_t t = idx[first + k];
            b.mx = vmax3(b.mx, tmx[t]);
            b.mn = vmin3(b.mn, tmn[t]);
        }
        return b;
    }

    // BVHNode::findBestSplitPlane (Geometry.h:293-365)
    float best_split(const TmpNode& nd, int& axis, float& pos_out) const {
        float parent_area = nd.b.area();
        float best = FLT_MAX;
        for (int ax = 0; ax < 3; ++ax) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            for (int32_t k = 0; k < nd.count; ++k) {
                float c = cen[idx[nd.first + k]][ax];
                lo = smin(lo, c);
                hi = smax(hi, c);
            }
            if (lo == hi) continue;
            Box bins[kBins];
            int cnt[kBins] = {0};
            float scale = kBins / (hi - lo);
            for (int32_t k = 0; k < nd.count; ++k) {
                int32_t t = idx[nd.first + k];
                int bi = std::min(kBins - 1, (int)((cen[t][ax] - lo) * scale));
                cnt[bi]++;
                bins[bi].mx = vmax3(bins[bi].mx, tmx[t]);
                bins[bi].mn = vmin3(bins[bi].mn, tmn[t]);
            }
            float la[kBins - 1], ra[kBins - 1];
            int lc[kBins - 1], rc[kBins - 1];
            int ls = 0, rs = 0;
            Box lb, rb;
            for (int i = 0; i < kBins - 1; ++i) {
                ls += cnt[i];
                lc[i] = ls;
                if (cnt[i] > 0) { lb.grow(bins[i].mn); lb.grow(bins[i].mx); }
                la[i] = (ls > 0) ? lb.area() : 0.f;
                rs += cnt[kBins - 1 - i];
                rc[kBins - 2 - i] = rs;
                if (cnt[kBins - 1 - i] > 0) { rb.grow(bins[kBins - 1 - i].mn); rb.grow(bins[kBins - 1 - i].mx); }
                ra[kBins - 2 - i] = (rs > 0) ? rb.area() : 0.f;
            }
            float step = (hi - lo) / kBins;
            for (int i = 0; i < kBins - 1; ++i) {
                float lr = la[i] / parent_area, rr = ra[i] / parent_area;
                float cost = (kBoundsCost + ((lr * (float)lc[i]) * kIsectCost)) + ((rr * (float)rc[i]) * kIsectCost);
                float cand = lo + step * (float)(i + 1);
                if (cost < best) { pos_out = cand; axis = ax; best = cost; }
            }
        }
        return best;
    }

    // BVHNode::subdivide (Geometry.h:367-417)
    void subdivide(int32_t ni, int par_depth) {
        TmpNode& nd = tmp[ni];
        uint32_t d = nd.depth;
        uint32_t cur = max_depth.load();
        while (d > cur && !max_depth.compare_exchange_weak(cur, d)) {}
        if (nd.count <= kLeafTris) return;
        int axis = -1;
        float split = -FLT_MAX;
        float sc = best_split(nd, axis, split);
        float no_split = (float)nd.count * kIsectCost;
        if (sc >= no_split || axis == -1) return;
        int32_t i = nd.first, j = i + nd.count - 1;
        while (i <= j) {
            if (cen[idx[i]][axis] < split) i++;
            else std::swap(idx[i], idx[j--]);
        }
        int32_t lcount = i - nd.first;
        if (lcount == 0 || lcount == nd.count) return;
        int32_t l = next.fetch_add(2);
        int32_t r = l + 1;
        tmp[l].first = nd.first; tmp[l].count = lcount; tmp[l].depth = d + 1;
        tmp[r].first = i; tmp[r].count = nd.count - lcount; tmp[r].depth = d + 1;
        nd.child = l;
        nd.count = 0;
        tmp[l].b = range_bounds(tmp[l].first, tmp[l].count);
        tmp[r].b = range_bounds(tmp[r].first, tmp[r].count);
        if (par_depth > 0 && tmp[r].count > 20000) {
            std::thread th([this, r, par_depth] { subdivide(r, par_depth - 1); });
            subdivide(l, par_depth - 1);
            th.join();
        } else {
            subdivide(l, 0);
            subdivide(r, 0);
        }
    }
};

void renumber(const std::vector<TmpNode>& tmp, int32_t ti, int32_t ri, std::vector<BvhNode>& out,
              int32_t& next) {
    // Iterative form of the reference's allocation order: a node's children get the
    // next pair of indices when it splits, then the left subtree is finished before
    // the right one starts.
    struct Item { int32_t t, r; };
    std::vector<Item> stack;
    stack.push_back({ti, ri});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const TmpNode& s = tmp[it.t];
        BvhNode& o = out[it.r];
        o.bmin[0] = s.b.mn.x; o.bmin[1] = s.b.mn.y; o.bmin[2] = s.b.mn.z;
        o.bmax[0] = s.b.mx.x; o.bmax[1] = s.b.mx.y; o.bmax[2] = s.b.mx.z;
        if (s.child >= 0) {
            int32_t l = next;
            next += 2;
            o.left_first = l;
            o.count = 0;
            stack.push_back({s.child + 1, l + 1});
            stack.push_back({s.child, l});
        } else {
            o.left_first = s.first;
            o.count = s.count;
        }
    }
}

}  // namespace

void build_bvh(const float* positions, uint32_t n_tris, HostBvh& out) {
    auto t0 = std::chrono::steady_clock::now();
    out.nodes.clear();
    out.tri_index.clear();
    out.max_depth = 0;
    if (n_tris == 0) { out.build_ms = 0.0; return; }
    Builder b;
    b.pos = positions;
    b.n = n_tris;
    b.cen.resize(n_tris); b.tmn.resize(n_tris); b.tmx.resize(n_tris); b.idx.resize(n_tris);
    for (uint32_t t = 0; t < n_tris; ++t) {
        HV3 p0 = b.vert(t, 0), p1 = b.vert(t, 1), p2 = b.vert(t, 2);
        b.cen[t] = ((p0 + p1) + p2) * 0.3333f;                                 // Geometry.h:96-98
        b.tmn[t] = vmin3(vmin3(vmin3(HV3(FLT_MAX, FLT_MAX, FLT_MAX), p0), p1), p2);
        b.tmx[t] = vmax3(vmax3(vmax3(HV3(-FLT_MAX, -FLT_MAX, -FLT_MAX), p0), p1), p2);
        b.idx[t] = (int32_t)t;
    }
    b.tmp.resize((size_t)2 * n_tris);
    b.tmp[0].first = 0; b.tmp[0].count = (int32_t)n_tris; b.tmp[0].depth = 0;
    b.tmp[0].b = b.range_bounds(0, (int32_t)n_tris);
    unsigned hw = std::thread::hardware_concurrency();
    int par = hw >= 16 ? 5 : (hw >= 4 ? 3 : 0);
    b.subdivide(0, par);
    int32_t used = b.next.load();
    out.nodes.assign((size_t)used, BvhNode{});
    int32_t nxt = 1;
    renumber(b.tmp, 0, 0, out.nodes, nxt);
    out.tri_index = std::move(b.idx);
    out.max_depth = b.max_depth.load();
    auto t1 = std::chrono::steady_clock::now();
    out.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// ------------------------------------------------------------------ camera
CameraRec make_camera(const pg_camera_desc& d) {
    // Matrix::perspective (MyMath.h:350-364) with near 0.001, far 10000 (SceneLoader.h:229)
    M4 P;
    std::memset(P.m, 0, sizeof(P.m));
    const float n = 0.001f, f = 10000.0f;
    float aspect = d.width / d.height;
    float t = 1.f / (tanf(d.fov_deg * 0.5f * 3.141592654f / 180.f));
    float inv = 1.f / (f - n);
    P.m[0] = t / aspect; P.m[5] = t; P.m[10] = -f * inv; P.m[11] = -(f * n) * inv; P.m[14] = -1.f;
    // Matrix::lookAt (MyMath.h:325-348), then inverted (SceneLoader.h:236-237)
    HV3 from(d.from[0], d.from[1], d.from[2]), to(d.to[0], d.to[1], d.to[2]), up(d.up[0], d.up[1], d.up[2]);
    HV3 dir = norm3(from - to);
    HV3 left = norm3(cross3(up, dir));
    HV3 nup = cross3(dir, left);
    M4 L;
    L.m[0] = left.x; L.m[1] = left.y; L.m[2] = left.z;
    L.m[4] = nup.x; L.m[5] = nup.y; L.m[6] = nup.z;
    L.m[8] = dir.x; L.m[9] = dir.y; L.m[10] = dir.z;
    L.m[3] = -dot3(from, left); L.m[7] = -dot3(from, nup); L.m[11] = -dot3(from, dir);
    L.m[15] = 1.f;
    M4 V = L.inverse();
    CameraRec c;
    M4 IP = P.inverse();                                                       // Scene.h:27
    std::memcpy(c.inv_proj, IP.m, 64);
    std::memcpy(c.cam, V.m, 64);
    HV3 o = V.mul_point(HV3(0.f, 0.f, 0.f));                                  // Scene.h:39
    c.origin[0] = o.x; c.origin[1] = o.y; c.origin[2] = o.z;
    c.width = d.width;
    c.height = d.height;
    return c;
}

// ------------------------------------------------------------------ environment
void build_env_tables(const float* rgb, int32_t w, int32_t h, EnvTables& out) {
    const double kPi = 3.14159265358979323846;
    out.w = w; out.h = h;
    size_t n = (size_t)w * h;
    out.texels.assign(n * 4, 0.f);
    std::vector<float> fn(n);
    std::vector<float> row_int(h);
    for (int v = 0; v < h; ++v) {
        float st = (float)sin(kPi * ((v + 0.5f) / h));                         // Sampling.h:178-189
        float c = 0.f;
        for (int u = 0; u < w; ++u) {
            size_t i = (size_t)v * w + u;
            float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
            out.texels[4 * i] = r; out.texels[4 * i + 1] = g; out.texels[4 * i + 2] = b;
            float lum = ((0.2126f * r) + (0.7152f * g)) + (0.0722f * b);
            fn[i] = lum * st;
        }
        for (int u = 0; u < w; ++u) c = c + (fn[(size_t)v * w + u] / w);       // Sampling.h:108-111
        row_int[v] = c;
    }
    float mc = 0.f;
    for (int v = 0; v < h; ++v) mc = mc + (row_int[v] / h);
    out.marginal_integral = mc;
    out.pdf.resize(n);
    for (size_t i = 0; i < n; ++i) out.pdf[i] = (mc == 0.f) ? 1.f : fn[i] / mc;  // Sampling.h:203-208
    // totalIntegratedPower (Lights.h:200-210)
    float total = 0.f;
    for (int i = 0; i < h; ++i) {
        float st = sinf((float)(((float)i / (float)h) * M_PI));
        for (int k = 0; k < w; ++k) {
            size_t j = (size_t)i * w + k;
            float lum = ((0.2126f * rgb[3 * j]) + (0.7152f * rgb[3 * j + 1])) + (0.0722f * rgb[3 * j + 2]);
            total += lum * st;
        }
    }
    total = total / (float)(w * h);
    out.power = (float)(total * 4.f * M_PI);
    // Vose alias table over texel weights pdf[i] (deterministic stack order).
    out.prob.assign(n, 1.f);
    out.alias.resize(n);
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += (double)out.pdf[i];
    std::vector<double> p(n);
    std::vector<uint32_t> small, large;
    small.reserve(n); large.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        p[i] = sum > 0.0 ? (double)out.pdf[i] * (double)n / sum : 1.0;
        out.alias[i] = (uint32_t)i;
        if (p[i] < 1.0) small.push_back((uint32_t)i); else large.push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t l = small.back(); small.pop_back();
        uint32_t g = large.back(); large.pop_back();
        out.prob[l] = (float)p[l];
        out.alias[l] = g;
        p[g] = (p[g] + p[l]) - 1.0;
        if (p[g] < 1.0) small.push_back(g); else large.push_back(g);
    }
    for (uint32_t g : large) { out.prob[g] = 1.f; out.alias[g] = g; }
    for (uint32_t l : small) { out.prob[l] = 1.f; out.alias[l] = l; }
}

// ------------------------------------------------------------------ packing
bool material_record(const pg_material& m, MaterialRec& r) {
    std::memset(&r, 0, sizeof(r));
    r.type = (int32_t)m.type;
    bool layered = r.type >= PG_BSDF_LAYERED;                                 // Materials.h:855-897
    if (layered) r.type -= PG_BSDF_LAYERED;
    if (r.type < PG_BSDF_DIFFUSE || r.type > PG_BSDF_PLASTIC) return false;
    r.tex = -1;
    for (int c = 0; c < 3; ++c) {
        r.albedo[c] = m.albedo[c]; r.emission[c] = m.emission[c]; r.eta[c] = m.eta[c]; r.k[c] = m.k[c];
    }
    r.int_ior = m.int_ior; r.ext_ior = m.ext_ior;
    r.alpha = 1.62142f * sqrtf(m.roughness);                                   // Materials.h:249, 455, 738
    if (r.type == PG_BSDF_ORENNAYAR) r.alpha = m.roughness;                   // sigma (Materials.h:649-652)
    float lum = ((0.2126f * m.emission[0]) + (0.7152f * m.emission[1])) + (0.0722f * m.emission[2]);
    r.is_light = lum > 0.f ? 1u : 0u;                                          // Materials.h:130
    bool spec = r.type == PG_BSDF_MIRROR || r.type == PG_BSDF_GLASS ||
                ((r.type == PG_BSDF_CONDUCTOR || r.type == PG_BSDF_DIELECTRIC) && r.alpha < kEps);
    r.specular = spec ? 1u : 0u;
    r.two_sided = (!layered && (r.type == PG_BSDF_GLASS || r.type == PG_BSDF_DIELECTRIC)) ? 0u : 1u;
    return true;
}

std::string prepare_scene(const float* pos, const float* nrm, const uint32_t* mat_idx, uint32_t n_tris,
                          const pg_material* mats, uint32_t n_mats, const float* env_rgb, int32_t env_w,
                          int32_t env_h, const pg_camera_desc& cam, PreparedScene& out, HostBvh* prebuilt) {
    if (n_tris == 0) return "scene has no triangles";
    // traversal stack entries hold node and triangle indices in 27 bits (2N-1 nodes)
    if (n_tris > (1u << 26)) return "scene has more than 2^26 triangles";
    if (!pos || !nrm || !mat_idx || !mats || n_mats == 0) return "null scene array";
    if (cam.width < 1.f || cam.height < 1.f) return "camera resolution must be >= 1x1";
    for (uint32_t i = 0; i < n_tris; ++i)
        if (mat_idx[i] >= n_mats) return "material index out of range";
    out.mats.resize(n_mats);
    for (uint32_t i = 0; i < n_mats; ++i)
        if (!material_record(mats[i], out.mats[i])) return "unknown material type";
    if (prebuilt) out.bvh = std::move(*prebuilt);  // built elsewhere (the GPU builder) in the same topology
    else build_bvh(pos, n_tris, out.bvh);
    out.node_slots.assign(out.bvh.nodes.size() + 2, BvhNode{});
    for (size_t i = 0; i < out.bvh.nodes.size(); ++i) out.node_slots[i + 1] = out.bvh.nodes[i];
    out.finite_bounds = true;
    for (const BvhNode& b : out.bvh.nodes)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(b.bmin[c]) || !std::isfinite(b.bmax[c])) out.finite_bounds = false;
    out.tris.resize(n_tris);
    for (uint32_t k = 0; k < n_tris; ++k) {
        uint32_t t = (uint32_t)out.bvh.tri_index[k];
        const float* p = pos + 9 * (size_t)t;
        TriIsect& o = out.tris[k];
        for (int c = 0; c < 3; ++c) {
            o.v0[c] = p[c];
            o.e0[c] = p[3 + c] - p[c];
            o.e1[c] = p[6 + c] - p[c];
        }
        o.id = t; o.pad0 = 0.f; o.pad1 = 0.f;
    }
    out.shade.resize(n_tris);
    Box bounds;
    for (uint32_t t = 0; t < n_tris; ++t) {
        const float* p = pos + 9 * (size_t)t;
        const float* q = nrm + 9 * (size_t)t;
        HV3 v0(p[0], p[1], p[2]), v1(p[3], p[4], p[5]), v2(p[6], p[7], p[8]);
        bounds.grow(v0); bounds.grow(v1); bounds.grow(v2);                    // Scene.h:122-127
        HV3 e1 = v2 - v1, e2 = v0 - v2;                                       // Geometry.h:86-93
        HV3 n = norm3(cross3(e1, e2));
        float area = len3(cross3(e1, e2)) * 0.5f;
        HV3 n0(q[0], q[1], q[2]);
        HV3 gn = n * (dot3(n0, n) > 0.f ? 1.f : -1.f);                        // Geometry.h:153
        TriShade& s = out.shade[t];
        for (int c = 0; c < 3; ++c) { s.n0[c] = q[c]; s.n1[c] = q[3 + c]; s.n2[c] = q[6 + c]; }
        s.gn[0] = gn.x; s.gn[1] = gn.y; s.gn[2] = gn.z;
        s.mat = mat_idx[t]; s.area = area; s.pad = 0.f; s.pad2 = 0.f;
    }
    out.has_env = env_rgb != nullptr && env_w > 0 && env_h > 0;
    out.lights.clear();
    if (out.has_env) {
        build_env_tables(env_rgb, env_w, env_h, out.env);
        if (out.env.power > 0.f) {                                             // Scene.h:133-136
            LightRec l;
            std::memset(&l, 0, sizeof(l));
            l.kind = 1; l.tri = -1;
            out.lights.push_back(l);
        }
    }
    for (uint32_t t = 0; t < n_tris; ++t) {                                    // Scene.h:93-100
        const MaterialRec& m = out.mats[mat_idx[t]];
        if (!m.is_light) continue;
        LightRec l;
        std::memset(&l, 0, sizeof(l));
        const float* p = pos + 9 * (size_t)t;
        for (int c = 0; c < 3; ++c) {
            l.p0[c] = p[c]; l.p1[c] = p[3 + c]; l.p2[c] = p[6 + c];
            l.emission[c] = m.emission[c]; l.gn[c] = out.shade[t].gn[c];
        }
        l.kind = 0; l.tri = (int32_t)t; l.area = out.shade[t].area;
        out.lights.push_back(l);
    }
    HV3 centre = (bounds.mx + bounds.mn) * 0.5f;                               // SceneLoader.h:267-268
    out.centre[0] = centre.x; out.centre[1] = centre.y; out.centre[2] = centre.z;
    out.radius = len3(bounds.mx - centre);
    out.camera = make_camera(cam);
    return std::string();
}

}  // namespace pg
