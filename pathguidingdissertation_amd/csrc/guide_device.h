// guide_device.h — device-side queries on the radiance cache: exact kNN over the
// point BVH and quadtree sampling / pdf evaluation.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "guide.h"

namespace pg {

// Squared distance from q to an AABB; per axis the gap is measured with the same
// rounded subtraction a contained point would use, so a point inside the box is
// never nearer than the box (pruning on `box > worst` is exact).
__device__ __forceinline__ float box_dist2(float4 lo, float4 hi, F3 q) {
    float dx = (q.x < lo.x) ? (lo.x - q.x) : ((q.x > hi.x) ? (q.x - hi.x) : 0.f);
    float dy = (q.y < lo.y) ? (lo.y - q.y) : ((q.y > hi.y) ? (q.y - hi.y) : 0.f);
    float dz = (q.z < lo.z) ? (lo.z - q.z) : ((q.z > hi.z) ? (q.z - hi.z) : 0.f);
    return ((dx * dx) + (dy * dy)) + (dz * dz);
}

__device__ __forceinline__ uint32_t orig_index(const GuideView& g, uint32_t s) {
    return __float_as_uint(g.rec_pos[s].w);
}

// (d, original index) lexicographic order: the k nearest set is unique.
__device__ __forceinline__ bool knn_better(const GuideView& g, float da, uint32_t sa, float db, uint32_t sb) {
    if (da < db) return true;
    if (da > db) return false;
    if (sb == 0xffffffffu) return sa != 0xffffffffu;  // empty slots sort last
    if (sa == 0xffffffffu) return false;
    return orig_index(g, sa) < orig_index(g, sb);
}

// K nearest records to q. bd/bs: squared distances and sorted record positions,
// ascending. stk: LDS stack (node, box distance), stride `stride`.
// cap: an upper bound on the K-th nearest squared distance (e.g. the largest distance
// among K known records, knn_cap), used for pruning until the candidate list is
// tighter; subtrees are pruned only when strictly farther, so the result is the same
// exact set (ties included) as without it.
template <int K>
__device__ void knn_query(const GuideView& g, F3 q, float (&bd)[K], uint32_t (&bs)[K], int* stk_node,
                          float* stk_dist, int stride, float cap = __int_as_float(0x7f800000)) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        bd[j] = __int_as_float(0x7f800000);  // +inf
        bs[j] = 0xffffffffu;
    }
    if (g.n_records == 0) return;
    const uint32_t P = g.n_leaf_pow2;
    int node = 0;
    int sp = 0;
    while (true) {
        if ((uint32_t)node >= P - 1) {
            uint32_t leaf = (uint32_t)node - (P - 1);
            uint32_t b = leaf * kLeafRecords;
            uint32_t e = min(b + (uint32_t)kLeafRecords, g.n_records);
            for (uint32_t s = b; s < e; ++s) {
                float4 p = g.rec_pos[s];
                float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                float d = ((dx * dx) + (dy * dy)) + (dz * dz);
                if (knn_better(g, d, s, bd[K - 1], bs[K - 1])) {
                    bd[K - 1] = d;
                    bs[K - 1] = s;
#pragma unroll
                    for (int j = K - 1; j > 0; --j) {
                        if (knn_better(g, bd[j], bs[j], bd[j - 1], bs[j - 1])) {
                            float td = bd[j]; bd[j] = bd[j - 1]; bd[j - 1] = td;
                            uint32_t ts = bs[j]; bs[j] = bs[j - 1]; bs[j - 1] = ts;
                        }
                    }
                }
            }
        } else {
            int c1 = 2 * node + 1, c2 = c1 + 1;
            float d1 = box_dist2(g.nodes[2 * c1], g.nodes[2 * c1 + 1], q);
            float d2 = box_dist2(g.nodes[2 * c2], g.nodes[2 * c2 + 1], q);
            float worst = fminf(bd[K - 1], cap);
            bool v1 = !(d1 > worst), v2 = !(d2 > worst);
            if (v1 && v2) {
                int nn = c1, ff = c2;
                float fd = d2;
                if (d2 < d1) { nn = c2; ff = c1; fd = d1; }
                stk_node[sp * stride] = ff;
                stk_dist[sp * stride] = fd;
                ++sp;
                node = nn;
                continue;
            }
            if (v1) { node = c1; continue; }
            if (v2) { node = c2; continue; }
        }
        // pop the next node that can still hold a candidate
        bool found = false;
        while (sp > 0) {
            --sp;
            float d = stk_dist[sp * stride];
            if (!(d > fminf(bd[K - 1], cap))) {
                node = stk_node[sp * stride];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
}

// Approximate neighbours for the primary-sample-space guide: the leaf that holds the
// query's 30-bit Morton key in the sorted record order (a binary search over rec_key,
// the key computed in the frame the records were sorted in) and that leaf's list of the
// kKnnMax records nearest its centroid, computed once per cache build (k_leaf_nbrs).
// No stack and no candidate heap per query.
__device__ __forceinline__ uint32_t morton_expand10(uint32_t x) {
    x &= 1023u;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ uint32_t morton_leaf(const GuideView& g, F3 q) {
    int qx = min(max((int)((q.x - g.m_lo[0]) * g.m_inv * 1024.f), 0), 1023);
    int qy = min(max((int)((q.y - g.m_lo[1]) * g.m_inv * 1024.f), 0), 1023);
    int qz = min(max((int)((q.z - g.m_lo[2]) * g.m_inv * 1024.f), 0), 1023);
    uint32_t key = (morton_expand10((uint32_t)qx) << 2) | (morton_expand10((uint32_t)qy) << 1) |
                   morton_expand10((uint32_t)qz);
    uint32_t lo = 0, hi = g.n_records;  // first sorted record with key >= query key
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (g.rec_key[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return min(lo, g.n_records - 1u) / (uint32_t)kLeafRecords;
}
__device__ __forceinline__ void nearby_records(const GuideView& g, F3 q, uint32_t* out) {
    for (int j = 0; j < kKnnMax; ++j) out[j] = 0xffffffffu;
    if (g.n_records == 0) return;
    const uint32_t* nb = g.leaf_nbrs + (size_t)morton_leaf(g, q) * kKnnMax;
    for (int j = 0; j < kKnnMax; ++j) out[j] = nb[j];
}

// Pruning bound for an exact k-nearest search (K = kKnnMax): the largest squared
// distance from q to the records of its leaf list (nearby_records). Those are K
// distinct records, so the K-th nearest distance is at most this value.
__device__ __forceinline__ float knn_cap(const GuideView& g, F3 q) {
    float cap = __int_as_float(0x7f800000);
    if (g.n_records < (uint32_t)kKnnMax) return cap;
    const uint32_t* nb = g.leaf_nbrs + (size_t)morton_leaf(g, q) * kKnnMax;
    float m = 0.f;
#pragma unroll 4
    for (int c = 0; c < kKnnMax; ++c) {
        uint32_t s = nb[c];
        if (s == 0xffffffffu) return cap;
        float4 p = g.rec_pos[s];
        float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        m = fmaxf(m, ((dx * dx) + (dy * dy)) + (dz * dz));
    }
    return m;
}

// Quadtree pdf over [0,1)^2 at (u, v).
__device__ __forceinline__ float qt_pdf_uv(const QtNode* nodes, float u, float v) {
    float pdf = 1.f;
    int node = 0;
    float ox = 0.f, oy = 0.f, size = 1.f;
    for (int depth = 0; depth < kQtMaxDepth + 1; ++depth) {
        QtNode nd = nodes[node];
        float tot = (nd.e[0] + nd.e[1]) + (nd.e[2] + nd.e[3]);
        float half = size * 0.5f;
        int qx = (u >= ox + half) ? 1 : 0, qy = (v >= oy + half) ? 1 : 0;
        int qd = qx + 2 * qy;
        if (tot > 0.f) pdf *= 4.f * nd.e[qd] / tot;
        int ch = nd.child[qd];
        if (ch < 0 || !(tot > 0.f)) break;
        ox += qx * half;
        oy += qy * half;
        size = half;
        node = ch;
    }
    return pdf;
}

// Sample (u, v) from a quadtree: the column (qx) from r1, the row (qy) from r2, both
// rescaled at each level; the remaining fractions place the point in the leaf.
__device__ __forceinline__ void qt_sample_uv(const QtNode* nodes, float r1, float r2, float& u, float& v,
                                             float& pdf) {
    pdf = 1.f;
    int node = 0;
    float ox = 0.f, oy = 0.f, size = 1.f;
    for (int depth = 0; depth < kQtMaxDepth + 1; ++depth) {
        QtNode nd = nodes[node];
        float tot = (nd.e[0] + nd.e[1]) + (nd.e[2] + nd.e[3]);
        float half = size * 0.5f;
        int qx, qy;
        if (tot > 0.f) {
            float left = nd.e[0] + nd.e[2];
            float pl = left / tot;
            if (r1 < pl) { qx = 0; r1 = r1 / pl; } else { qx = 1; r1 = (r1 - pl) / (1.f - pl); }
            float col = nd.e[qx] + nd.e[qx + 2];
            float pb = col > 0.f ? nd.e[qx] / col : 0.5f;
            if (r2 < pb) { qy = 0; r2 = r2 / pb; } else { qy = 1; r2 = (r2 - pb) / (1.f - pb); }
            pdf *= 4.f * nd.e[qx + 2 * qy] / tot;
        } else {
            qx = r1 < 0.5f ? 0 : 1; r1 = qx ? (r1 - 0.5f) * 2.f : r1 * 2.f;
            qy = r2 < 0.5f ? 0 : 1; r2 = qy ? (r2 - 0.5f) * 2.f : r2 * 2.f;
        }
        r1 = fminf(fmaxf(r1, 0.f), 0.99999994f);
        r2 = fminf(fmaxf(r2, 0.f), 0.99999994f);
        int ch = nd.child[qx + 2 * qy];
        ox += qx * half;
        oy += qy * half;
        size = half;
        if (ch < 0 || !(tot > 0.f)) break;
        node = ch;
    }
    u = ox + r1 * size;
    v = oy + r2 * size;
}

}  // namespace pg
