// device_shading.h — device functions shared by the wavefront kernels: BVH
// traversal (Geometry.h:448-567), camera rays (Scene.h:45-57), environment map
// (Lights.h:128-228), BSDFs (Materials.h:36-639) and shadow-ray setup (Scene.h:139-146).
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "kernels.h"

namespace pg {

constexpr int kBlock = 128;

__device__ __forceinline__ int f2i(float f) { return __float_as_int(f); }

// AABB::rayAABB (Geometry.h:175-192)
__device__ __forceinline__ bool slab(float4 lo, float4 hi, F3 o, F3 inv, float& t) {
    float t0x = (lo.x - o.x) * inv.x, t0y = (lo.y - o.y) * inv.y, t0z = (lo.z - o.z) * inv.z;
    float t1x = (hi.x - o.x) * inv.x, t1y = (hi.y - o.y) * inv.y, t1z = (hi.z - o.z) * inv.z;
    float ex = smin(t0x, t1x), ey = smin(t0y, t1y), ez = smin(t0z, t1z);
    float xx = smax(t0x, t1x), xy = smax(t0y, t1y), xz = smax(t0z, t1z);
    float fe = smax(ex, smax(ey, ez));
    float fx = smin(xx, smin(xy, xz));
    if (fe > fx || fx < 0.f) return false;
    t = smax(0.f, fe);
    return true;
}

// slab() with IEEE min/max (v_min/v_max, three-operand forms): the same hit decision
// and entry distance whenever no product is NaN, i.e. with finite box bounds, origin
// and 1/d; the ternary min/max of slab() differ from IEEE min/max only on NaN operands
// and on +0/-0 ties, and no comparison made with the result tells the zeros apart.
__device__ __forceinline__ bool slab_ieee(float4 lo, float4 hi, F3 o, F3 inv, float& t) {
    float t0x = (lo.x - o.x) * inv.x, t0y = (lo.y - o.y) * inv.y, t0z = (lo.z - o.z) * inv.z;
    float t1x = (hi.x - o.x) * inv.x, t1y = (hi.y - o.y) * inv.y, t1z = (hi.z - o.z) * inv.z;
    float fe = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    float fx = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    if (fe > fx || fx < 0.f) return false;
    t = fmaxf(0.f, fe);
    return true;
}

// Triangle::rayIntersect (Geometry.h:101-130) on a leaf-ordered record already
// loaded as its three float4 rows (v0 + id, e0, e1).
__device__ __forceinline__ bool mt_rows(float4 a, float4 b, float4 c, F3 o, F3 d, float& t, float& u, float& v) {
    F3 v0{a.x, a.y, a.z}, e0{b.x, b.y, b.z}, e1{c.x, c.y, c.z};
    F3 p = cross(d, e1);
    float det = dot(e0, p);
    if (fabsf(det) < kMtEps) return false;
    float inv_det = 1.f / det;
    F3 T = o - v0;
    u = dot(T, p) * inv_det;
    if (u < 0.f || u > 1.f) return false;
    F3 qv = cross(T, e0);
    v = dot(d, qv) * inv_det;
    if (v < 0.f || v > 1.f || (u + v > 1.f)) return false;
    t = dot(e1, qv) * inv_det;
    return t >= 0.f;
}

__device__ __forceinline__ bool mt(const TriIsect* tp, F3 o, F3 d, float& t, float& u, float& v) {
    const float4* q = reinterpret_cast<const float4*>(tp);
    return mt_rows(q[0], q[1], q[2], o, d, t, u, v);
}

__device__ __forceinline__ void node_header(const BvhNode* nodes, int idx, int& lf, int& cnt) {
    const int* h = reinterpret_cast<const int*>(nodes + idx + 1);
    lf = h[3];
    cnt = h[7];
}

// Traversal stack entry: the pushed node's (leftFirst, count) so that a pop costs no
// memory access. Bits 31..5 hold leftFirst and bits 4..0 the count; a leaf of 31 or
// more triangles is stored as its node index with code 31, and its pop reads the
// header. Node and triangle indices must stay below 2^27 (checked at scene upload).
constexpr uint32_t kStkEscape = 31u;
__device__ __forceinline__ int stk_pack(int node, int lf, int cnt) {
    return (int)((uint32_t)cnt < kStkEscape ? ((uint32_t)lf << 5) | (uint32_t)cnt
                                            : ((uint32_t)node << 5) | kStkEscape);
}
__device__ __forceinline__ void stk_pop(const BvhNode* nodes, int e, int& lf, int& cnt) {
    const uint32_t u = (uint32_t)e, c = u & 31u;
    if (c == kStkEscape) {
        node_header(nodes, (int)(u >> 5), lf, cnt);
    } else {
        lf = (int)(u >> 5);
        cnt = (int)c;
    }
}

// A lane's traversal stack with every entry in LDS: entry k at p[k * kBlock].
struct LdsStack {
    int* p;
    __device__ __forceinline__ void push(int& sp, int e) const { p[(sp++) * kBlock] = e; }
    __device__ __forceinline__ int pop(int& sp) const { return p[(--sp) * kBlock]; }
};
// A lane's traversal stack whose first NL entries live in LDS and deeper ones (rarely
// reached) in global memory: entry k >= NL at spill[(k - NL) * stride]. A smaller LDS
// share per lane lets more waves run on a CU; the entries and their order are those of
// the all-LDS stack, so every traversal visits the same nodes.
// The two parts carry their address spaces in their pointer types (LDS: 3, global: 1)
// so that a pop stays a ds_read plus a rare global load instead of one generic load.
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(1))) int global_int;
template <int NL>
struct SplitStack {
    lds_int* lds;
    global_int* spill;
    uint32_t stride;
    __device__ __forceinline__ void push(int& sp, int e) const {
        if (sp < NL) lds[sp * kBlock] = e;
        else spill[(size_t)(sp - NL) * stride] = e;
        ++sp;
    }
    __device__ __forceinline__ int pop(int& sp) const {
        --sp;
        int e = lds[(sp < NL ? sp : NL - 1) * kBlock];
        if (sp >= NL) e = spill[(size_t)(sp - NL) * stride];
        return e;
    }
};

// Resumable closest-hit traversal of one lane: ray, best hit so far, current node
// (leftFirst, count), next triangle of a leaf and stack depth. trav_begin tests the
// root; each trav_step advances BVHNode::traverse's loop (Geometry.h:448-512) by one
// triangle or one interior node.
struct TravState {
    F3 o, d, inv;
    float ht, hu, hv;
    uint32_t hid;
    int lf, cnt, sp;
    int k;      // next triangle of the current leaf (0 at an interior node)
    bool ieee;  // box tests may use slab_ieee (no NaN can arise for this ray)
};

// Box tests of this ray may use IEEE min/max: finite scene bounds, origin and 1/d.
__device__ __forceinline__ bool ieee_boxes(const DevScene& s, const TravState& t) {
    return s.finite_bounds && isfinite(t.inv.x) && isfinite(t.inv.y) && isfinite(t.inv.z) && isfinite(t.o.x) &&
           isfinite(t.o.y) && isfinite(t.o.z);
}

// Entry distances of an interior node's two children, kFltMax for a child that is
// missed or starts beyond the current hit (Geometry.h:486-497).
__device__ __forceinline__ void child_dists(const TravState& t, float4 a0, float4 a1, float4 b0, float4 b1, float& dl,
                                            float& dr) {
    dl = 0.f;
    dr = 0.f;
    if (t.ieee) {
        if (!slab_ieee(a0, a1, t.o, t.inv, dl) || dl >= t.ht) dl = kFltMax;
        if (!slab_ieee(b0, b1, t.o, t.inv, dr) || dr >= t.ht) dr = kFltMax;
    } else {
        if (!slab(a0, a1, t.o, t.inv, dl) || dl >= t.ht) dl = kFltMax;
        if (!slab(b0, b1, t.o, t.inv, dr) || dr >= t.ht) dr = kFltMax;
    }
}

template <bool COUNT>
__device__ __forceinline__ bool trav_begin(const DevScene& s, TravState& t, uint32_t& cn) {
    t.inv = F3{1.f / t.d.x, 1.f / t.d.y, 1.f / t.d.z};
    t.ieee = ieee_boxes(s, t);
    const float4* nf = reinterpret_cast<const float4*>(s.nodes);
    float4 r0 = nf[2], r1 = nf[3];  // root at slot 1
    float tb;
    if (COUNT) cn += 1;
    if (!slab(r0, r1, t.o, t.inv, tb)) return false;
    if (tb > t.ht) return false;
    t.lf = f2i(r0.w);
    t.cnt = f2i(r1.w);
    t.sp = 0;
    t.k = 0;
    return true;
}

// One step of one lane: a single triangle of the current leaf, or an interior node's
// two children (near first, far pushed); a leaf's last triangle or a node whose
// children are both missed pops the stack. Every lane of a wave issues the same four
// 16-byte loads per step (the node pair, or a 48-byte triangle record and 16 bytes
// after it, hence the padded triangle array), whichever kind of step it takes; the
// per-lane order of tests is that of BVHNode::traverse (Geometry.h:448-512): triangles
// in leaf order (ties keep the first, as :475), children in entry order. False when
// the traversal has finished.
template <bool COUNT, class Stack>
__device__ __forceinline__ bool trav_step(const DevScene& s, TravState& t, const Stack& stk, uint32_t& cn,
                                          uint32_t& ct) {
    const bool leaf = t.cnt > 0;
    const float4* q = leaf ? reinterpret_cast<const float4*>(s.tris + t.lf + t.k)
                           : reinterpret_cast<const float4*>(s.nodes) + 2 * (t.lf + 1);
    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    bool pop;
    if (leaf) {
        float tt, u, v;
        if (COUNT) ct += 1;
        if (mt_rows(q0, q1, q2, t.o, t.d, tt, u, v) && tt > kEps && tt < t.ht) {
            t.ht = tt;
            t.hid = __float_as_uint(q0.w);
            t.hu = u;
            t.hv = v;
        }
        pop = ++t.k == t.cnt;
    } else {
        if (COUNT) cn += 2;
        float dl, dr;
        child_dists(t, q0, q1, q2, q3, dl, dr);
        int nr = t.lf + 1;
        int lfl = f2i(q0.w), cl = f2i(q1.w), lfr = f2i(q2.w), cr = f2i(q3.w);
        if (dl > dr) {
            float tt = dl; dl = dr; dr = tt;
            nr = t.lf;
            int ti = lfl; lfl = lfr; lfr = ti;
            ti = cl; cl = cr; cr = ti;
        }
        pop = dl == kFltMax;
        if (!pop) {
            t.lf = lfl;
            t.cnt = cl;
            t.k = 0;
            if (dr != kFltMax) stk.push(t.sp, stk_pack(nr, lfr, cr));
        }
    }
    if (pop) {
        if (t.sp == 0) return false;
        stk_pop(s.nodes, stk.pop(t.sp), t.lf, t.cnt);
        t.k = 0;
    }
    return true;
}

// Closest hit. stk: this lane's traversal stack (LdsStack / SplitStack).
template <bool COUNT, class Stack>
__device__ __forceinline__ void traverse_closest(const DevScene& s, F3 o, F3 d, uint32_t& hid, float& ht,
                                                 float& hu, float& hv, const Stack& stk, uint32_t& cn, uint32_t& ct) {
    TravState t;
    t.o = o;
    t.d = d;
    t.ht = ht;
    t.hid = hid;
    t.hu = hu;
    t.hv = hv;
    if (trav_begin<COUNT>(s, t, cn))
        while (trav_step<COUNT>(s, t, stk, cn, ct)) {
        }
    hid = t.hid;
    ht = t.ht;
    hu = t.hu;
    hv = t.hv;
}

// Resumable any-hit traversal (BVHNode::traverseVisible, Geometry.h:514-567) with the
// TravState fields (ht holds maxT); vis_step returns 0 while running, 1 when the ray
// is unoccluded, 2 when a triangle occludes it.
template <bool COUNT>
__device__ __forceinline__ bool vis_begin(const DevScene& s, TravState& t, uint32_t& cn) {
    t.inv = F3{1.f / t.d.x, 1.f / t.d.y, 1.f / t.d.z};
    t.ieee = ieee_boxes(s, t);
    const float4* nf = reinterpret_cast<const float4*>(s.nodes);
    float4 r0 = nf[2], r1 = nf[3];
    float tb;
    if (COUNT) cn += 1;
    if (!slab(r0, r1, t.o, t.inv, tb)) return false;
    if (tb > t.ht) return false;
    t.lf = f2i(r0.w);
    t.cnt = f2i(r1.w);
    t.sp = 0;
    t.k = 0;
    return true;
}

// One step of an any-hit traversal, as trav_step (same loads for every lane):
// 0 while running, 1 when the ray is unoccluded, 2 when a triangle occludes it.
template <bool COUNT, class Stack>
__device__ __forceinline__ int vis_step(const DevScene& s, TravState& t, const Stack& stk, uint32_t& cn, uint32_t& ct) {
    const bool leaf = t.cnt > 0;
    const float4* q = leaf ? reinterpret_cast<const float4*>(s.tris + t.lf + t.k)
                           : reinterpret_cast<const float4*>(s.nodes) + 2 * (t.lf + 1);
    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    bool pop;
    if (leaf) {
        float tt, u, v;
        if (COUNT) ct += 1;
        if (mt_rows(q0, q1, q2, t.o, t.d, tt, u, v) && tt > kEps && tt < t.ht) return 2;
        pop = ++t.k == t.cnt;
    } else {
        if (COUNT) cn += 2;
        float dl, dr;
        child_dists(t, q0, q1, q2, q3, dl, dr);
        int nr = t.lf + 1;
        int lfl = f2i(q0.w), cl = f2i(q1.w), lfr = f2i(q2.w), cr = f2i(q3.w);
        if (dl > dr) {
            float tt = dl; dl = dr; dr = tt;
            nr = t.lf;
            int ti = lfl; lfl = lfr; lfr = ti;
            ti = cl; cl = cr; cr = ti;
        }
        pop = dl == kFltMax;
        if (!pop) {
            t.lf = lfl;
            t.cnt = cl;
            t.k = 0;
            if (dr != kFltMax) stk.push(t.sp, stk_pack(nr, lfr, cr));
        }
    }
    if (pop) {
        if (t.sp == 0) return 1;
        stk_pop(s.nodes, stk.pop(t.sp), t.lf, t.cnt);
        t.k = 0;
    }
    return 0;
}

// Any hit in (1e-4, max_t): returns true when unoccluded.
template <bool COUNT, class Stack>
__device__ __forceinline__ bool traverse_visible(const DevScene& s, F3 o, F3 d, float max_t, const Stack& stk,
                                                 uint32_t& cn, uint32_t& ct) {
    TravState t;
    t.o = o;
    t.d = d;
    t.ht = max_t;
    if (!vis_begin<COUNT>(s, t, cn)) return true;
    int r;
    while ((r = vis_step<COUNT>(s, t, stk, cn, ct)) == 0) {
    }
    return r == 1;
}

// ------------------------------------------------------------------ camera
// Camera::generateRay (Scene.h:45-57) with Matrix::mulPointAndPerspectiveDivide / mulVec.
__device__ __forceinline__ void generate_ray(const CameraRec& c, float x, float y, F3& o, F3& d) {
    float xc = (2.f * x / c.width) - 1.f;
    float yc = (2.f * (1.f - (y / c.height))) - 1.f;
    const float* m = c.inv_proj;
    float zc = 0.f;
    F3 p{((xc * m[0] + yc * m[1]) + zc * m[2]) + m[3], ((xc * m[4] + yc * m[5]) + zc * m[6]) + m[7],
         ((xc * m[8] + yc * m[9]) + zc * m[10]) + m[11]};
    float w = ((m[12] * xc + m[13] * yc) + m[14] * zc) + m[15];
    w = 1.f / w;
    F3 dc = p * w;
    const float* k = c.cam;
    F3 dw{(dc.x * k[0] + dc.y * k[1]) + dc.z * k[2], (dc.x * k[4] + dc.y * k[5]) + dc.z * k[6],
          (dc.x * k[8] + dc.y * k[9]) + dc.z * k[10]};
    d = normalize(dw);
    o = F3{c.origin[0], c.origin[1], c.origin[2]};
}

// ------------------------------------------------------------------ environment (Lights.h:128-228)
constexpr double kPiD = 3.14159265358979323846;

__device__ __forceinline__ C3 env_lookup(const DevScene& s, float tu, float tv) {  // Imaging.h:82-108
    float u = smax(0.f, fabsf(tu)) * (float)s.env_w;
    float v = smax(0.f, fabsf(tv)) * (float)s.env_h;
    int x = (int)floorf(u), y = (int)floorf(v);
    float fu = u - (float)x, fv = v - (float)y;
    float w0 = (1.f - fu) * (1.f - fv), w1 = fu * (1.f - fv), w2 = (1.f - fu) * fv, w3 = fu * fv;
    x = x % s.env_w;
    y = y % s.env_h;
    int x1 = (x + 1) % s.env_w, y1 = (y + 1) % s.env_h;
    float4 a = s.env_tex[y * s.env_w + x], b = s.env_tex[y * s.env_w + x1];
    float4 c = s.env_tex[y1 * s.env_w + x], e = s.env_tex[y1 * s.env_w + x1];
    C3 r = ((c3(a.x, a.y, a.z) * w0 + c3(b.x, b.y, b.z) * w1) + c3(c.x, c.y, c.z) * w2) + c3(e.x, e.y, e.z) * w3;
    return r;
}
__device__ __forceinline__ void env_uv(F3 wi, float& u, float& v) {
    u = atan2f(wi.z, wi.x);
    u = (u < 0.f) ? (float)((double)u + 2.0 * kPiD) : u;
    u = (float)((double)u / (2.0 * kPiD));
    v = (float)((double)acosf(wi.y) / kPiD);
}
__device__ __forceinline__ C3 env_eval(const DevScene& s, F3 wi) {
    float u, v;
    env_uv(wi, u, v);
    return env_lookup(s, u, v);
}
__device__ __forceinline__ float env_pdf(const DevScene& s, F3 wi) {
    float u, v;
    env_uv(wi, u, v);
    float st = (float)sin(kPiD * (double)v);
    if (st <= 0.f) return 0.f;
    int iu = min(max((int)(u * (float)s.env_w), 0), s.env_w - 1);
    int iv = min(max((int)(v * (float)s.env_h), 0), s.env_h - 1);
    return (float)((double)s.env_pdf[iv * s.env_w + iu] / (2.0 * kPiD * kPiD * (double)st));
}
// Alias-table sample: 4 draws (texel, acceptance, in-texel u, in-texel v).
__device__ __forceinline__ F3 env_sample(const DevScene& s, uint32_t key, uint32_t& dim, C3& col, float& pdf) {
    int n = s.env_w * s.env_h;
    float r1 = keyed_rand(key, dim++), r2 = keyed_rand(key, dim++);
    float r3 = keyed_rand(key, dim++), r4 = keyed_rand(key, dim++);
    int i = min((int)(r1 * (float)n), n - 1);
    if (!(r2 < s.env_prob[i])) i = (int)s.env_alias[i];
    int iu = i % s.env_w, iv = i / s.env_w;
    float u = ((float)iu + r3) / (float)s.env_w;
    float v = ((float)iv + r4) / (float)s.env_h;
    float theta = (float)(kPiD * (double)v);
    float phi = (float)(2.0 * kPiD * (double)u);
    float ct = cosf(theta), cp = cosf(phi), st = sinf(theta), sp = sinf(phi);
    F3 wi{cp * st, ct, sp * st};
    pdf = (st <= 0.f) ? 0.f : (float)((double)s.env_pdf[i] / (2.0 * kPiD * kPiD * (double)st));
    col = (pdf <= 0.f) ? c3(0.f, 0.f, 0.f) : env_eval(s, wi);
    return wi;
}

// ------------------------------------------------------------------ BSDFs (Materials.h)
constexpr float kInvPiF = 0.318309886183790671538f;

__device__ __forceinline__ float fresnel_dielectric(float cos_theta, float iint, float iext) {  // :36-61
    float ci = smax(smin(cos_theta, 1.f), -1.f);
    float eta = (ci < 0.f) ? (iint / iext) : (iext / iint);
    if (ci < 0.f) ci = fabsf(ci);
    float eta2 = eta * eta;
    float ci2 = smax(ci * ci, 0.f);
    float si2 = smax(1.f - ci2, 0.f);
    float st2 = eta2 * si2;
    if (st2 > 1.f) return 1.f;
    float ctt = sqrtf(smax(1.f - st2, 0.f));
    float par = (ci - eta * ctt) / (ci + eta * ctt);
    float per = (eta * ci - ctt) / (eta * ci + ctt);
    return ((par * par) + (per * per)) * 0.5f;
}
__device__ __forceinline__ C3 fresnel_conductor(float cos_theta, C3 ior, C3 k) {  // :63-90
    float ci = smax(smin(cos_theta, 1.f), -1.f);
    float ci2 = smax(ci * ci, 0.f);
    float si2 = smax(1.f - ci2, 0.f);
    C3 cC = c3(ci, ci, ci), cC2 = c3(ci2, ci2, ci2), cS2 = c3(si2, si2, si2);
    C3 ior2 = ior * ior, k2 = k * k;
    C3 par = (((ior2 + k2) * cC2) - (ior * cC * 2.f) + cS2) / (((ior2 + k2) * cC2) + (ior * cC * 2.f) + cS2);
    C3 per = ((ior2 + k2) - (ior * cC * 2.f) + cC2) / ((ior2 + k2) + (ior * cC * 2.f) + cC2);
    return (par + per) * 0.5f;
}
__device__ __forceinline__ float lambda_ggx(F3 w, float alpha) {  // :93-100
    if (w.z <= 0.f) return 0.f;
    float a2 = alpha * alpha, c2 = w.z * w.z;
    float s2 = smax(1.f - c2, 0.f);
    float t2 = fabsf(s2 / c2);
    return (sqrtf(1.f + a2 * t2) - 1.f) * 0.5f;
}
__device__ __forceinline__ float g_ggx(F3 wi, F3 wo, float alpha) {
    float g1o = 1.f / (1.f + lambda_ggx(wo, alpha));
    float g1i = 1.f / (1.f + lambda_ggx(wi, alpha));
    return g1o * g1i;
}
__device__ __forceinline__ float d_ggx(F3 h, float alpha) {  // :111-118
    if (h.z <= 0.f) return 0.f;
    float a2 = alpha * alpha, c2 = h.z * h.z;
    float q = c2 * (a2 - 1.f) + 1.f;
    float den = (float)((kPiD * (double)q) * (double)q);
    if (den <= 0.f) return 0.f;
    return a2 / den;
}
__device__ __forceinline__ float cosine_pdf(F3 wl) {  // Sampling.h:98-101
    float theta = acosf(wl.z);
    float c = cosf(theta);
    return (c > 0.f) ? (float)((double)c * 0.318309886183790671538) : 0.f;
}

struct SD {
    F3 x, wo, sn, gn;
    Frame fr;
    C3 alb;  // the material's albedo (its texture at the hit's uv when it has one)
};

// Albedo texture lookup (Imaging.h:82-108) at the hit's interpolated uv
// (Triangle::interpolateAttributes, Geometry.h:135-136).
__device__ __forceinline__ C3 tex_albedo(const DevScene& s, int tex, uint32_t tri, float a, float b, float g) {
    const float2* uv = s.uv + 3 * (size_t)tri;
    float2 u0 = uv[0], u1 = uv[1], u2 = uv[2];
    float tu = (u0.x * a + u1.x * b) + u2.x * g;
    float tv = (u0.y * a + u1.y * b) + u2.y * g;
    int4 ti = s.tex_info[tex];
    float u = smax(0.f, fabsf(tu)) * (float)ti.y;
    float v = smax(0.f, fabsf(tv)) * (float)ti.z;
    int x = (int)floorf(u), y = (int)floorf(v);
    float fu = u - (float)x, fv = v - (float)y;
    float w0 = (1.f - fu) * (1.f - fv), w1 = fu * (1.f - fv), w2 = (1.f - fu) * fv, w3 = fu * fv;
    x = x % ti.y;
    y = y % ti.z;
    int x1 = (x + 1) % ti.y, y1 = (y + 1) % ti.z;
    const float4* t = s.tex_texels + ti.x;
    float4 p = t[y * ti.y + x], q = t[y * ti.y + x1], r = t[y1 * ti.y + x], e = t[y1 * ti.y + x1];
    return ((c3(p.x, p.y, p.z) * w0 + c3(q.x, q.y, q.z) * w1) + c3(r.x, r.y, r.z) * w2) + c3(e.x, e.y, e.z) * w3;
}
__device__ __forceinline__ C3 material_albedo(const DevScene& s, const MaterialRec& m, uint32_t tri, float a,
                                              float b, float g) {
    return m.tex >= 0 ? tex_albedo(s, m.tex, tri, a, b, g) : c3(m.albedo[0], m.albedo[1], m.albedo[2]);
}

__device__ __forceinline__ float spherical_phi(F3 w) {  // MyMath.h:445-448
    float p = atan2f(w.y, w.x);
    return (p < 0.f) ? (float)((double)p + 2.0 * kPiD) : p;
}
// OrenNayarBSDF::evaluate (Materials.h:683-707): no hemisphere test, as in the reference.
__device__ __forceinline__ C3 oren_nayar_eval(const MaterialRec& m, const SD& sd, F3 wl, F3 ol) {
    float s2 = m.alpha * m.alpha;  // alpha holds sigma for this type
    float A = 1.f - (s2 / (2.f * (s2 + 0.33f)));
    float B = (0.45f * s2) / (s2 + 0.09f);
    float phi_i = spherical_phi(wl), phi_o = spherical_phi(ol);
    float th_i = acosf(wl.z), th_o = acosf(ol.z);
    float c = cosf(phi_i - phi_o);
    float sa = sinf(smax(th_i, th_o));
    float tb = tanf(smin(th_i, th_o));
    float on = A + (((B * smax(0.f, c)) * sa) * tb);
    return (sd.alb * kInvPiF) * on;
}
__device__ __forceinline__ float plastic_exponent(float alpha) {  // Materials.h:738-740
    float a = smax(alpha, 0.001f);
    return (2.f / (a * a)) - 2.f;
}
// PlasticBSDF::evaluate / PDF (Materials.h:784-848): FP64 where the reference's M_PI
// terms promote.
__device__ __forceinline__ C3 plastic_eval(const MaterialRec& m, const SD& sd, F3 wl, F3 ol, float& pdf) {
    pdf = 0.f;
    if (wl.z <= 0.f || ol.z <= 0.f) return c3(0.f, 0.f, 0.f);
    F3 h = wl + ol;
    if (length_sq(h) < kEps) return c3(0.f, 0.f, 0.f);
    h = normalize(h);
    float e = plastic_exponent(m.alpha);
    float oh = dot(ol, h);
    float ks = fresnel_dielectric(oh, m.int_ior, m.ext_ior);
    float kd = 1.f - ks;
    float norm = (float)((double)(e + 2.f) / ((4.0 * kPiD) * (double)(2.f - powf(2.f, (-e * 0.5f)))));
    float lo = (float)((double)(e + 2.f) / (8.0 * kPiD)), hi = (float)((double)(e + 4.f) / (8.0 * kPiD));
    norm = smin(smax(lo, norm), hi);
    float pw = powf(smax(h.z, 0.f), e);
    float ph = (float)(((double)(e + 1.f) / (2.0 * kPiD)) * (double)pw);
    float gp = oh <= 0.f ? 0.f : (ph / (4.f * oh));
    pdf = (kd * cosine_pdf(wl)) + (ks * gp);
    C3 diff = sd.alb * kInvPiF;
    float gl = oh <= 0.f ? 0.f : pw;
    return (diff * kd) + ((c3(gl, gl, gl) * norm) * ks);
}

__device__ inline C3 bsdf_eval(const MaterialRec& m, const SD& sd, F3 wi, float& pdf) {
    pdf = 0.f;
    C3 zero = c3(0.f, 0.f, 0.f);
    C3 alb = sd.alb;
    if (m.type == PG_BSDF_ORENNAYAR) {  // Materials.h:683-714
        F3 wl = to_local(sd.fr, wi);
        pdf = cosine_pdf(wl);
        return oren_nayar_eval(m, sd, wl, to_local(sd.fr, sd.wo));
    }
    if (m.type == PG_BSDF_PLASTIC) return plastic_eval(m, sd, to_local(sd.fr, wi), to_local(sd.fr, sd.wo), pdf);
    if (m.type == PG_BSDF_DIFFUSE) {
        F3 wl = to_local(sd.fr, wi);
        pdf = cosine_pdf(wl);
        if (wl.z <= 0.f) return zero;
        return alb * kInvPiF;
    }
    if (m.type == PG_BSDF_CONDUCTOR) {
        F3 wl = to_local(sd.fr, wi), ol = to_local(sd.fr, sd.wo);
        if (wl.z <= 0.f || ol.z <= 0.f) return zero;
        if (m.alpha < kEps) return zero;
        F3 wm = wl + ol;
        if (length_sq(wm) < kEps) return zero;
        wm = normalize(wm);
        if (dot(ol, wm) <= 0.f) return zero;
        if (wl.z * ol.z <= 0.f) return zero;
        float D = d_ggx(wm, m.alpha);
        pdf = (D * wm.z) / (4.f * dot(ol, wm));
        float G = g_ggx(wl, ol, m.alpha);
        C3 F = fresnel_conductor(dot(ol, wm), c3(m.eta[0], m.eta[1], m.eta[2]), c3(m.k[0], m.k[1], m.k[2]));
        return alb * (((F * G) * D) / ((4.f * wl.z) * ol.z));
    }
    if (m.type == PG_BSDF_DIELECTRIC) {
        F3 ol = to_local(sd.fr, sd.wo), wl = to_local(sd.fr, wi);
        if (ol.z == 0.f || wl.z == 0.f) return zero;
        float eta = ol.z > 0.f ? m.ext_ior / m.int_ior : m.int_ior / m.ext_ior;
        if (eta == 1.f || m.alpha < kEps) return zero;
        bool refl = ol.z * wl.z > 0.f;
        float ep = 1.f;
        if (!refl) ep = ol.z > 0.f ? eta : 1.f / eta;
        F3 wm = wl * ep + ol;
        if (length_sq(wm) < kEps) return zero;
        wm = normalize(wm);
        if (dot(wm, wl) * wl.z < 0.f || dot(wm, ol) * ol.z < 0.f) return zero;
        float F = fresnel_dielectric(dot(ol, wm), m.int_ior, m.ext_ior);
        if (refl) {
            if (dot(ol, wm) <= 0.f) return zero;
            if (wl.z * ol.z == 0.f) return zero;
            float D = d_ggx(wm, m.alpha);
            pdf = F * ((D * wm.z) / (4.f * dot(ol, wm)));
            float G = g_ggx(wl, ol, m.alpha);
            return alb * (((F * G) * D) / ((4.f * wl.z) * ol.z));
        }
        float sm = dot(wl, wm) + dot(ol, wm) / ep;
        float den = sm * sm;
        if (den <= 0.f) return zero;
        float D = d_ggx(wm, m.alpha);
        pdf = (1.f - F) * ((D * fabsf(dot(ol, wm))) / den);
        float G = g_ggx(wl, ol, m.alpha);
        return ((alb * ((((1.f - F) * G) * D) / den)) * (fabsf(dot(wl, wm)) / fabsf(wl.z))) *
               (fabsf(dot(ol, wm)) / fabsf(ol.z));
    }
    return zero;  // mirror / glass: delta distributions
}

// The random numbers come from `rng.next()` in the order the reference's sample()
// draws them: the counter RNG in the renderer, given numbers in pg_bsdf_query.
template <class Rng>
__device__ inline F3 bsdf_sample(const MaterialRec& m, const SD& sd, Rng& rng, C3& f, float& pdf) {
    C3 alb = sd.alb;
    F3 up{0.f, 0.f, 1.f};
    f = c3(0.f, 0.f, 0.f);
    pdf = 0.f;
    switch (m.type) {
    case PG_BSDF_PLASTIC: {  // Materials.h:742-778
        F3 ol = to_local(sd.fr, sd.wo);
        float r1 = rng.next(), r2 = rng.next();
        float F = fresnel_dielectric(fabsf(ol.z), m.int_ior, m.ext_ior);
        F3 wl;
        if (rng.next() < F) {
            float e = plastic_exponent(m.alpha);
            float base = 1.f - r1, power = 1.f / (e + 1.f);
            float th = acosf(powf(base, power));
            float ph = (float)(2.0 * kPiD * (double)r2);
            F3 h = normalize(spherical_to_world(th, ph));
            wl = ((h * 2.f) * dot(ol, h)) - ol;
        } else {
            wl = spherical_to_world(acosf(sqrtf(r1)), (float)(2.0 * kPiD * (double)r2));
        }
        F3 wi = to_world(sd.fr, wl);
        f = bsdf_eval(m, sd, wi, pdf);
        return wi;
    }
    case PG_BSDF_ORENNAYAR:  // Materials.h:655-670: the diffuse cosine warp
    case PG_BSDF_DIFFUSE: {  // Materials.h:146-161
        float r1 = rng.next(), r2 = rng.next();
        float theta = acosf(sqrtf(r1));
        float phi = (float)(2.0 * kPiD * (double)r2);
        F3 wi = to_world(sd.fr, spherical_to_world(theta, phi));
        f = bsdf_eval(m, sd, wi, pdf);
        return wi;
    }
    case PG_BSDF_MIRROR: {  // Materials.h:201-215
        F3 ol = to_local(sd.fr, sd.wo);
        F3 wr{-ol.x, -ol.y, ol.z};
        if (fabsf(wr.z) <= kEps) return up;
        f = alb / wr.z;
        pdf = 1.f;
        return to_world(sd.fr, wr);
    }
    case PG_BSDF_CONDUCTOR: {  // Materials.h:253-285
        F3 ol = to_local(sd.fr, sd.wo);
        if (ol.z <= 0.f) return up;
        if (m.alpha < kEps) {
            F3 wr{-ol.x, -ol.y, ol.z};
            C3 F = fresnel_conductor(wr.z, c3(m.eta[0], m.eta[1], m.eta[2]), c3(m.k[0], m.k[1], m.k[2]));
            f = (alb / fabsf(wr.z)) * F;
            pdf = 1.f;
            return to_world(sd.fr, wr);
        }
        float a2 = m.alpha * m.alpha;
        float r1 = rng.next(), r2 = rng.next();
        float tm = acosf(sqrtf((1.f - r1) / (r1 * (a2 - 1.f) + 1.f)));
        float pm = (float)(2.0 * kPiD * (double)r2);
        F3 wm = spherical_to_world(tm, pm);
        F3 wl = -ol + (wm * 2.f) * dot(wm, ol);
        F3 wi = to_world(sd.fr, wl);
        f = bsdf_eval(m, sd, wi, pdf);
        return wi;
    }
    case PG_BSDF_GLASS: {  // Materials.h:375-421
        F3 ol = to_local(sd.fr, sd.wo);
        if (fabsf(ol.z) < kEps) return up;
        float F = fresnel_dielectric(ol.z, m.int_ior, m.ext_ior);
        float ci = ol.z;
        float eta = (ci < 0.f) ? (m.int_ior / m.ext_ior) : (m.ext_ior / m.int_ior);
        if (ci < 0.f) ci = fabsf(ci);
        if (rng.next() < F) {
            F3 wr{-ol.x, -ol.y, ol.z};
            pdf = F;
            f = (alb / fabsf(wr.z)) * F;
            return to_world(sd.fr, wr);
        }
        float ci2 = smax(ci * ci, 0.f), si2 = smax(1.f - ci2, 0.f);
        float eta2 = eta * eta, st2 = eta2 * si2;
        if (st2 > 1.f) {
            F3 wr{-ol.x, -ol.y, ol.z};
            pdf = F;
            f = (alb / fabsf(wr.z)) * F;
            return to_world(sd.fr, wr);
        }
        float ctt = sqrtf(smax(1.f - st2, 0.f));
        float wz = (ol.z < 0.f) ? ctt : -ctt;
        F3 wt{-eta * ol.x, -eta * ol.y, wz};
        pdf = 1.f - F;
        f = ((alb / fabsf(wt.z)) * (1.f - F)) * eta2;
        return to_world(sd.fr, wt);
    }
    default: {  // PG_BSDF_DIELECTRIC, Materials.h:459-539
        F3 ol = to_local(sd.fr, sd.wo);
        if (m.alpha < kEps) {
            float F = fresnel_dielectric(ol.z, m.int_ior, m.ext_ior);
            if (rng.next() < F) {
                F3 wl{-ol.x, -ol.y, ol.z};
                pdf = F;
                f = (alb / fabsf(wl.z)) * F;
                return to_world(sd.fr, wl);
            }
            float eta = ol.z > 0.f ? m.ext_ior / m.int_ior : m.int_ior / m.ext_ior;
            float eta2 = eta * eta;
            float sg = ol.z > 0.f ? -1.f : 1.f;
            float ci2 = smax(ol.z * ol.z, 0.f), si2 = smax(1.f - ci2, 0.f), st2 = eta2 * si2;
            if (st2 > 1.f) {
                F3 wr{-ol.x, -ol.y, ol.z};
                pdf = F;
                f = (alb / fabsf(wr.z)) * F;
                return to_world(sd.fr, wr);
            }
            float ctt = sqrtf(smax(1.f - st2, 0.f));
            F3 wl{-eta * ol.x, -eta * ol.y, ctt * sg};
            pdf = 1.f - F;
            f = (alb / fabsf(wl.z)) * (1.f - F);
            return to_world(sd.fr, wl);
        }
        float a2 = m.alpha * m.alpha;
        float r1 = rng.next(), r2 = rng.next();
        float tm = acosf(sqrtf((1.f - r1) / (r1 * (a2 - 1.f) + 1.f)));
        float pm = (float)(2.0 * kPiD * (double)r2);
        F3 wm = spherical_to_world(tm, pm);
        float F = fresnel_dielectric(dot(ol, wm), m.int_ior, m.ext_ior);
        if (rng.next() < F) {
            F3 wl = -ol + (wm * 2.f) * dot(wm, ol);
            F3 wi = to_world(sd.fr, wl);
            f = bsdf_eval(m, sd, wi, pdf);
            return wi;
        }
        Frame mf = frame_from(wm);
        F3 orl = to_local(mf, ol);
        float ci = orl.z;
        float eta = ci > 0.f ? m.ext_ior / m.int_ior : m.int_ior / m.ext_ior;
        float eta2 = eta * eta;
        float sg = ci > 0.f ? -1.f : 1.f;
        float ci2 = smax(ci * ci, 0.f), si2 = smax(1.f - ci2, 0.f), st2 = eta2 * si2;
        float ctt = sqrtf(smax(1.f - st2, 0.f));
        F3 wl{-eta * orl.x, -eta * orl.y, ctt * sg};
        F3 wi = to_world(sd.fr, to_world(mf, wl));
        f = bsdf_eval(m, sd, wi, pdf);
        return wi;
    }
    }
}

struct KeyRng {  // the counter RNG stream of a path: keyed_rand(key, dim++)
    uint32_t key, dim;
    __device__ float next() { return keyed_rand(key, dim++); }
};
struct ListRng {  // given numbers, then 0.5 (the reference's test sampler)
    float u[3];
    int c;
    __device__ float next() { return c < 3 ? u[c++] : 0.5f; }
};
__device__ inline F3 bsdf_sample(const MaterialRec& m, const SD& sd, uint32_t key, uint32_t& dim, C3& f, float& pdf) {
    KeyRng r{key, dim};
    F3 wi = bsdf_sample(m, sd, r, f, pdf);
    dim = r.dim;
    return wi;
}

// BSDF::invert (Materials.h:163-171, 287-306, 672-681): the primary-sample-space
// point (u, v) that sample() maps to wi. Defined for the BSDFs whose sampling is a
// two-number warp: Diffuse and Oren-Nayar (cosine hemisphere) and the rough
// Conductor (GGX half vector); false for the rest (the reference's Plastic and
// Dielectric inverts are empty; delta BSDFs have none) and for directions the warp
// cannot produce (below the surface, or a half vector below it). At alpha = 1 the
// conductor's expression is 0/0; its limit u = 1 - cos^2(theta_m) is used there.
__device__ __forceinline__ bool pss_invertible(const MaterialRec& m) {
    return m.type == PG_BSDF_DIFFUSE || m.type == PG_BSDF_ORENNAYAR || (m.type == PG_BSDF_CONDUCTOR && !(m.alpha < kEps));
}
__device__ __forceinline__ bool bsdf_invert(const MaterialRec& m, const SD& sd, F3 wi, float& u, float& v) {
    F3 wl = to_local(sd.fr, wi);
    if (m.type == PG_BSDF_DIFFUSE || m.type == PG_BSDF_ORENNAYAR) {
        if (!(wl.z > 0.f)) return false;
        float c = cosf(acosf(wl.z));
        u = c * c;
        v = (float)((double)spherical_phi(wl) / (2.0 * kPiD));
    } else if (m.type == PG_BSDF_CONDUCTOR && !(m.alpha < kEps)) {
        F3 wm = normalize(wl + to_local(sd.fr, sd.wo));
        if (!(wm.z > 0.f)) return false;
        float c = cosf(acosf(wm.z));
        float a2 = m.alpha * m.alpha, am1 = a2 - 1.f;
        u = am1 == 0.f ? 1.f - c * c : (a2 / ((c * c) * (am1 * am1) + am1)) - (1.f / am1);
        v = (float)((double)spherical_phi(wm) / (2.0 * kPiD));
    } else {
        return false;
    }
    u = fminf(fmaxf(u, 0.f), 0.99999994f);
    v = fminf(fmaxf(v, 0.f), 0.99999994f);
    return true;
}
// sample() of an invertible BSDF with its two random numbers given (the warp that
// bsdf_invert inverts): Materials.h:146-161 / 655-670 (cosine) and 253-285 (GGX).
__device__ __forceinline__ F3 bsdf_warp(const MaterialRec& m, const SD& sd, float r1, float r2) {
    if (m.type == PG_BSDF_CONDUCTOR) {
        F3 ol = to_local(sd.fr, sd.wo);
        float a2 = m.alpha * m.alpha;
        float tm = acosf(sqrtf((1.f - r1) / (r1 * (a2 - 1.f) + 1.f)));
        F3 wm = spherical_to_world(tm, (float)(2.0 * kPiD * (double)r2));
        return to_world(sd.fr, -ol + (wm * 2.f) * dot(wm, ol));
    }
    return to_world(sd.fr, spherical_to_world(acosf(sqrtf(r1)), (float)(2.0 * kPiD * (double)r2)));
}

// Scene::visible (Scene.h:139-146): shadow ray from p1 toward p2.
__device__ __forceinline__ void make_visible_ray(F3 p1, F3 p2, F3& o, F3& d, float& max_t) {
    F3 dir = p2 - p1;
    max_t = length(dir) - (2.f * kEps);
    dir = normalize(dir);
    o = p1 + dir * kEps;
    d = dir;
}


}  // namespace pg
