// bvh_build.hip — the binned-SAH BVH build on the GPU (gfx950), level by level,
// producing the reference's exact topology, node numbering and triangle order
// (BVHNode::build / subdivide / findBestSplitPlane / updateBounds, Geometry.h:
// 274-446; 16 bins, leaf <= 4, costs 1 and 1.5).
//
// Every node of a level is handled at once by position-parallel kernels over the
// triangle-order array:
//   1. centroid bounds and 16-bin counts / bounds of each active node (atomics on
//      order-preserving integer images of the floats, privatised in LDS when a whole
//      block lies in one node);
//   2. one thread per node evaluates the 3 x 15 SAH candidates with the reference's
//      arithmetic and order (strict < keeps the first best);
//   3. the reference's in-place swap partition is reproduced exactly from two
//      device-wide prefix sums: with nL left elements, the right elements of the
//      prefix [0, nL) (ascending, table P, then P[k] = nL) trade places with the left
//      elements of the suffix (descending), and suffix position j receives
//      orig[j + 1] when that element goes right, else orig[P[#left after j]];
//   4. children's bounds: the first (in triangle order) of the elements equal to the
//      minimum / maximum, as the reference's sequential AABB::extend keeps, through
//      64-bit atomicMin on (order key, position).
// Nodes are allocated breadth-first; a last pass renumbers them depth-first the way
// the reference allocates child pairs (left subtree before right).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bvh_build.h"
#include "guiding.h"

namespace pg {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 1024;
constexpr int kBinWords = 3 * kBins * 7;  // per node: 3 axes x 16 bins x (count, min.xyz, max.xyz)

struct GNode {
    int32_t first, count, child, depth;
    float bmin[3], bmax[3];
};

__device__ __forceinline__ float fmin_ref(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }  // std::max

// Order-preserving image of a float; -0 and +0 map to the same key.
__device__ __forceinline__ uint32_t okey(float f) {
    uint32_t u = __float_as_uint(f == 0.f ? 0.f : f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ unsigned long long shfl_min64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t shfl_min32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ uint32_t shfl_max32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
// true when every active lane of the wave holds the same value
__device__ __forceinline__ bool wave_uniform(int v) {
    int first = __shfl(v, __ffsll((unsigned long long)__ballot(1)) - 1);
    return __all(v == first);
}

// Triangle::centre (Geometry.h:96-98) and the triangle's own bounds (the fold of
// AABB::extend over its three vertices, Geometry.h:274-287).
__global__ void k_tri_prep(const float* pos, uint32_t n, float4* cen, float4* tmn, float4* tmx, int32_t* idx_a,
                           int32_t* idx_b, int32_t* slot) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float* p = pos + 9 * (size_t)t;
    float c[3], lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        float a = p[k], b = p[3 + k], d = p[6 + k];
        c[k] = ((a + b) + d) * 0.3333f;
        lo[k] = fmin_ref(fmin_ref(fmin_ref(kFltMax, a), b), d);
        hi[k] = fmax_ref(fmax_ref(fmax_ref(-kFltMax, a), b), d);
    }
    cen[t] = make_float4(c[0], c[1], c[2], 0.f);
    tmn[t] = make_float4(lo[0], lo[1], lo[2], 0.f);
    tmx[t] = make_float4(hi[0], hi[1], hi[2], 0.f);
    idx_a[t] = (int32_t)t;
    idx_b[t] = (int32_t)t;
    slot[t] = n > (uint32_t)kLeafTris ? 0 : -1;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Bounds keys of target nodes: per node 6 x u64 (min x,y,z then max x,y,z), each
// (order key of the value, position in triangle order); the smallest key is the
// first position holding the extreme value.
__global__ void k_bound_keys(const int32_t* target, const int32_t* idx, const float4* tmn, const float4* tmx,
                             uint32_t n, unsigned long long* keys) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    int tn = p < n ? target[p] : -1;
    bool on = tn >= 0;
    unsigned long long k[6];
    if (on) {
        int t = idx[p];
        float4 a = tmn[t], b = tmx[t];
        k[0] = ((unsigned long long)okey(a.x) << 32) | p;
        k[1] = ((unsigned long long)okey(a.y) << 32) | p;
        k[2] = ((unsigned long long)okey(a.z) << 32) | p;
        k[3] = ((unsigned long long)(~okey(b.x)) << 32) | p;
        k[4] = ((unsigned long long)(~okey(b.y)) << 32) | p;
        k[5] = ((unsigned long long)(~okey(b.z)) << 32) | p;
    } else {
        for (int j = 0; j < 6; ++j) k[j] = ~0ull;
    }
    // one atomic per wave when the whole wave feeds one node
    if (__all(on) && wave_uniform(tn)) {
        for (int j = 0; j < 6; ++j) k[j] = shfl_min64(k[j]);
        if ((threadIdx.x & 63) == 0)
            for (int j = 0; j < 6; ++j) atomicMin(&keys[6 * (size_t)tn + j], k[j]);
        return;
    }
    if (on)
        for (int j = 0; j < 6; ++j) atomicMin(&keys[6 * (size_t)tn + j], k[j]);
}

// Node bounds from the keys: the value bits of the element at the winning position.
__global__ void k_bound_finalize(const unsigned long long* keys, const int32_t* idx, const float4* tmn,
                                 const float4* tmx, uint32_t n, int32_t node_begin, int32_t node_end, GNode* nodes) {
    int32_t nd = node_begin + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (nd >= node_end) return;
    GNode& g = nodes[nd];
    for (int j = 0; j < 3; ++j) {
        uint32_t pm = (uint32_t)(keys[6 * (size_t)nd + j] & 0xffffffffull);
        uint32_t px = (uint32_t)(keys[6 * (size_t)nd + 3 + j] & 0xffffffffull);
        if (pm >= n || px >= n) continue;  // every node holds >= 1 triangle; kept in bounds regardless
        float4 a = tmn[idx[pm]], b = tmx[idx[px]];
        g.bmin[j] = j == 0 ? a.x : (j == 1 ? a.y : a.z);
        g.bmax[j] = j == 0 ? b.x : (j == 1 ? b.y : b.z);
    }
}

// Step 1a: centroid bounds per active node (values only: the split search uses them
// in arithmetic where -0 and +0 act alike).
__global__ void k_centroid_bounds(const int32_t* slot, const int32_t* idx, const float4* cen, uint32_t n,
                                  uint32_t* cb) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    int a = p < n ? slot[p] : -1;
    bool on = a >= 0;
    uint32_t lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0u, 0u, 0u};
    if (on) {
        float4 c = cen[idx[p]];
        lo[0] = hi[0] = okey(c.x);
        lo[1] = hi[1] = okey(c.y);
        lo[2] = hi[2] = okey(c.z);
    }
    if (__all(on) && wave_uniform(a)) {
        for (int k = 0; k < 3; ++k) { lo[k] = shfl_min32(lo[k]); hi[k] = shfl_max32(hi[k]); }
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < 3; ++k) { atomicMin(&cb[6 * a + k], lo[k]); atomicMax(&cb[6 * a + 3 + k], hi[k]); }
        return;
    }
    if (on)
        for (int k = 0; k < 3; ++k) { atomicMin(&cb[6 * a + k], lo[k]); atomicMax(&cb[6 * a + 3 + k], hi[k]); }
}

// Bin word layout per node and axis: [bin][0] count, [1..3] min.xyz, [4..6] max.xyz.
__device__ __forceinline__ void bin_add(uint32_t* b, int ax, int bi, float4 lo, float4 hi) {
    uint32_t* w = b + (ax * kBins + bi) * 7;
    atomicAdd(&w[0], 1u);
    atomicMin(&w[1], okey(lo.x)); atomicMin(&w[2], okey(lo.y)); atomicMin(&w[3], okey(lo.z));
    atomicMax(&w[4], okey(hi.x)); atomicMax(&w[5], okey(hi.y)); atomicMax(&w[6], okey(hi.z));
}

// Step 1b: 16-bin counts and bounds per axis (Geometry.h:306-324).
__global__ __launch_bounds__(kBlock) void k_bins(const int32_t* slot, const int32_t* idx, const float4* cen,
                                                 const float4* tmn, const float4* tmx, uint32_t n, const uint32_t* cb,
                                                 uint32_t* bins) {
    __shared__ uint32_t lds[kBinWords];
    __shared__ int block_slot;
    const uint32_t b0 = blockIdx.x * blockDim.x;
    const uint32_t p = b0 + threadIdx.x;
    if (threadIdx.x == 0) {
        uint32_t last = min(b0 + blockDim.x, n) - 1;
        int s0 = slot[b0], s1 = slot[last];
        block_slot = (s0 >= 0 && s0 == s1) ? s0 : -1;  // node ranges are contiguous
    }
    __syncthreads();
    const int bs = block_slot;
    if (bs >= 0) {
        for (int k = threadIdx.x; k < kBinWords; k += blockDim.x) {
            int f = k % 7;
            lds[k] = f == 0 ? 0u : (f <= 3 ? ~0u : 0u);
        }
        __syncthreads();
    }
    int a = p < n ? slot[p] : -1;
    if (a >= 0) {
        int t = idx[p];
        float4 c = cen[t], lo = tmn[t], hi = tmx[t];
        uint32_t* dst = bs >= 0 ? lds : bins + (size_t)a * kBinWords;
        for (int ax = 0; ax < 3; ++ax) {
            float mn = okey_inv(cb[6 * a + ax]), mx = okey_inv(cb[6 * a + 3 + ax]);
            if (mn == mx) continue;
            float scale = (float)kBins / (mx - mn);
            float cv = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
            int bi = min(kBins - 1, (int)((cv - mn) * scale));
            bin_add(dst, ax, bi, lo, hi);
        }
    }
    if (bs >= 0) {
        __syncthreads();
        uint32_t* g = bins + (size_t)bs * kBinWords;
        for (int k = threadIdx.x; k < kBinWords; k += blockDim.x) {
            int f = k % 7;
            uint32_t v = lds[k];
            if (f == 0) { if (v) atomicAdd(&g[k], v); }
            else if (f <= 3) { if (v != ~0u) atomicMin(&g[k], v); }
            else if (v) atomicMax(&g[k], v);
        }
    }
}

struct Box3 {
    float mn[3], mx[3];
    __device__ void reset() {
        for (int k = 0; k < 3; ++k) { mn[k] = kFltMax; mx[k] = -kFltMax; }
    }
    __device__ void grow(const float* p) {
        for (int k = 0; k < 3; ++k) { mx[k] = fmax_ref(mx[k], p[k]); mn[k] = fmin_ref(mn[k], p[k]); }
    }
    // AABB::area (Geometry.h:212-215)
    __device__ float area() const {
        float sx = mx[0] - mn[0], sy = mx[1] - mn[1], sz = mx[2] - mn[2];
        return (((sx * sy) + (sy * sz)) + (sx * sz)) * 2.f;
    }
};

// Step 2: BVHNode::findBestSplitPlane + the subdivide decision (Geometry.h:293-382),
// one thread per active node. split_axis -1: the node stays a leaf.
__global__ void k_sah(const int32_t* act, uint32_t n_act, const GNode* nodes, const uint32_t* cb, const uint32_t* bins,
                      int32_t* split_axis, float* split_pos) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_act) return;
    const GNode& nd = nodes[act[a]];
    Box3 pb;
    for (int k = 0; k < 3; ++k) { pb.mn[k] = nd.bmin[k]; pb.mx[k] = nd.bmax[k]; }
    const float parent_area = pb.area();
    float best = kFltMax, pos_out = -kFltMax;
    int axis = -1;
    const uint32_t* B = bins + (size_t)a * kBinWords;
    for (int ax = 0; ax < 3; ++ax) {
        float lo = okey_inv(cb[6 * a + ax]), hi = okey_inv(cb[6 * a + 3 + ax]);
        if (lo == hi) continue;
        float la[kBins - 1], ra[kBins - 1];
        int lc[kBins - 1], rc[kBins - 1];
        int ls = 0, rs = 0;
        Box3 lb, rb;
        lb.reset();
        rb.reset();
        for (int i = 0; i < kBins - 1; ++i) {
            const uint32_t* wl = B + (ax * kBins + i) * 7;
            int cl = (int)wl[0];
            ls += cl;
            lc[i] = ls;
            if (cl > 0) {
                float mn[3] = {okey_inv(wl[1]), okey_inv(wl[2]), okey_inv(wl[3])};
                float mx[3] = {okey_inv(wl[4]), okey_inv(wl[5]), okey_inv(wl[6])};
                lb.grow(mn);
                lb.grow(mx);
            }
            la[i] = (ls > 0) ? lb.area() : 0.f;
            const uint32_t* wr = B + (ax * kBins + (kBins - 1 - i)) * 7;
            int cr = (int)wr[0];
            rs += cr;
            rc[kBins - 2 - i] = rs;
            if (cr > 0) {
                float mn[3] = {okey_inv(wr[1]), okey_inv(wr[2]), okey_inv(wr[3])};
                float mx[3] = {okey_inv(wr[4]), okey_inv(wr[5]), okey_inv(wr[6])};
                rb.grow(mn);
                rb.grow(mx);
            }
            ra[kBins - 2 - i] = (rs > 0) ? rb.area() : 0.f;
        }
        float step = (hi - lo) / (float)kBins;
        for (int i = 0; i < kBins - 1; ++i) {
            float lr = la[i] / parent_area, rr = ra[i] / parent_area;
            float cost = (kBoundsCost + ((lr * (float)lc[i]) * kIsectCost)) + ((rr * (float)rc[i]) * kIsectCost);
            float cand = lo + step * (float)(i + 1);
            if (cost < best) { pos_out = cand; axis = ax; best = cost; }
        }
    }
    float no_split = (float)nd.count * kIsectCost;
    if (best >= no_split || axis == -1) axis = -1;
    split_axis[a] = axis;
    split_pos[a] = pos_out;
}

// Step 3a: left / right flags of the positions of splitting nodes.
__global__ void k_flags(const int32_t* slot, const int32_t* idx, const float4* cen, uint32_t n,
                        const int32_t* split_axis, const float* split_pos, uint32_t* L, uint32_t* R) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = slot[p];
    uint32_t l = 0, r = 0;
    if (a >= 0 && split_axis[a] >= 0) {
        float4 c = cen[idx[p]];
        int ax = split_axis[a];
        float cv = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
        l = cv < split_pos[a] ? 1u : 0u;  // Geometry.h:389
        r = 1u - l;
    }
    L[p] = l;
    R[p] = r;
}

// exclusive prefix at p, valid for p == n too
__device__ __forceinline__ uint32_t excl(const uint32_t* S, const uint32_t* X, uint32_t n, uint32_t p) {
    return p < n ? S[p] : S[n - 1] + X[n - 1];
}

// Step 3b: per splitting node, nL and the last entry of the rank table P.
__global__ void k_split_counts(const int32_t* act, uint32_t n_act, const GNode* nodes, const int32_t* split_axis,
                               const uint32_t* L, const uint32_t* SL, const uint32_t* R, const uint32_t* SR,
                               uint32_t n, int32_t* n_left, int32_t* ptab) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_act) return;
    if (split_axis[a] < 0) { n_left[a] = -1; return; }
    const GNode& nd = nodes[act[a]];
    uint32_t f = (uint32_t)nd.first, e = f + (uint32_t)nd.count;
    int32_t nl = (int32_t)(excl(SL, L, n, e) - SL[f]);
    int32_t k = (int32_t)(excl(SR, R, n, f + (uint32_t)nl) - SR[f]);  // right elements of the prefix
    n_left[a] = nl;
    ptab[f + (uint32_t)k] = nl;
}

// Step 3c: P[m] = relative position of the m-th right element of the prefix.
__global__ void k_ptab(const int32_t* slot, uint32_t n, const int32_t* act, const GNode* nodes, const int32_t* n_left,
                       const uint32_t* L, const uint32_t* SR, int32_t* ptab) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = slot[p];
    if (a < 0 || n_left[a] < 0) return;
    uint32_t f = (uint32_t)nodes[act[a]].first;
    int32_t r = (int32_t)(p - f);
    if (r < n_left[a] && !L[p]) ptab[f + (SR[p] - SR[f])] = r;  // rank < nl: inside the node
}

// Step 3d: the partitioned order (BVHNode::subdivide's swap loop, Geometry.h:384-391).
__global__ void k_partition(const int32_t* slot, uint32_t n, const int32_t* act, const GNode* nodes,
                            const int32_t* n_left, const uint32_t* L, const uint32_t* SL, const int32_t* ptab,
                            const int32_t* in, int32_t* out) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = slot[p];
    if (a < 0) return;
    const int32_t nl = n_left[a];
    if (nl < 0) {  // no split attempted: the range is unchanged
        out[p] = in[p];
        return;
    }
    const GNode& nd = nodes[act[a]];
    const uint32_t f = (uint32_t)nd.first;
    const int32_t cnt = nd.count;
    const int32_t r = (int32_t)(p - f);
    const int32_t left_upto = (int32_t)(SL[p] - SL[f] + L[p]);  // left elements in [0, r]
    const int32_t after = nl - left_upto;                       // left elements in (r, cnt)
    if (r < nl) {
        if (L[p]) out[p] = in[p];
        return;  // right elements of the prefix receive the suffix's left elements below
    }
    const int32_t q = ptab[f + (uint32_t)after];  // after <= k <= nl: written by k_split_counts / k_ptab
    if (q < 0 || q >= cnt) return;                // never taken; keeps every access inside the node
    if (L[p]) out[f + (uint32_t)q] = in[p];
    if (r < cnt - 1 && !L[p + 1]) out[p] = in[p + 1];
    else out[p] = in[f + (uint32_t)q];
}

// Step 3e: children (a pair of nodes, breadth-first ids) of the nodes that split.
__global__ void k_children(const int32_t* act, uint32_t n_act, GNode* nodes, const int32_t* n_left,
                           uint32_t* counters, int32_t* child_slot, int32_t* next_act) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_act) return;
    int32_t nl = n_left[a];
    GNode& nd = nodes[act[a]];
    if (nl <= 0 || nl >= nd.count) return;  // no split, or the split failed (Geometry.h:393-395)
    int32_t l = (int32_t)atomicAdd(&counters[0], 2u);
    nd.child = l;
    GNode c0, c1;
    c0.first = nd.first; c0.count = nl; c0.child = -1; c0.depth = nd.depth + 1;
    c1.first = nd.first + nl; c1.count = nd.count - nl; c1.child = -1; c1.depth = nd.depth + 1;
    nodes[l] = c0;
    nodes[l + 1] = c1;
    atomicMax(&counters[2], (uint32_t)(nd.depth + 1));
    for (int k = 0; k < 2; ++k) {
        int32_t c = l + k;
        int32_t s = -1;
        if ((k ? c1.count : c0.count) > kLeafTris) {
            s = (int32_t)atomicAdd(&counters[1], 1u);
            next_act[s] = c;
        }
        child_slot[c] = s;
    }
}

// Step 4a: per position of a splitting node, its child (bounds target) and its
// active slot in the next level.
// A range that is final from this level on (a leaf) gets its order in both buffers,
// so the last level's output holds the whole triangle order.
__global__ void k_child_map(int32_t* slot, uint32_t n, const int32_t* act, const GNode* nodes, const int32_t* n_left,
                            const int32_t* child_slot, int32_t* target, int32_t* in, const int32_t* out) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = slot[p];
    target[p] = -1;
    if (a < 0) return;
    const GNode& nd = nodes[act[a]];
    if (nd.child < 0) {
        slot[p] = -1;  // the node stays a leaf
        in[p] = out[p];
        return;
    }
    int32_t c = nd.child + ((int32_t)(p - (uint32_t)nd.first) < n_left[a] ? 0 : 1);
    target[p] = c;
    slot[p] = child_slot[c];
    if (child_slot[c] < 0) in[p] = out[p];
}

// Renumbering: internal nodes per subtree (bottom up), then the reference's
// depth-first ids: a node that splits takes the next pair for its children, and the
// left subtree is numbered before the right one.
__global__ void k_internal(GNode* nodes, int32_t begin, int32_t end, int32_t* internal) {
    int32_t nd = begin + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (nd >= end) return;
    int32_t c = nodes[nd].child;
    internal[nd] = c < 0 ? 0 : 1 + internal[c] + internal[c + 1];
}
__global__ void k_dfs(const GNode* nodes, int32_t begin, int32_t end, const int32_t* internal, int32_t* dfs_id,
                      int32_t* dfs_base) {
    int32_t nd = begin + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (nd >= end) return;
    int32_t c = nodes[nd].child;
    if (c < 0) return;
    int32_t base = dfs_base[nd];
    dfs_id[c] = base;
    dfs_id[c + 1] = base + 1;
    dfs_base[c] = base + 2;
    dfs_base[c + 1] = base + 2 + 2 * internal[c];
}
__global__ void k_write_nodes(const GNode* nodes, int32_t n_nodes, const int32_t* dfs_id, const int32_t* dfs_base,
                              BvhNode* out) {
    int32_t nd = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (nd >= n_nodes) return;
    const GNode& g = nodes[nd];
    BvhNode o;
    for (int k = 0; k < 3; ++k) { o.bmin[k] = g.bmin[k]; o.bmax[k] = g.bmax[k]; }
    if (g.child >= 0) { o.left_first = dfs_base[nd]; o.count = 0; }
    else { o.left_first = g.first; o.count = g.count; }
    out[dfs_id[nd]] = o;
}

// Per-level accumulators: centroid minima at the largest key, maxima at 0; bin
// counts 0, bin minima at the largest key, maxima at 0.
__global__ void k_level_init(uint32_t* cb, uint32_t* bins, uint32_t n_act) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)n_act * 6) cb[i] = (i % 6) < 3 ? ~0u : 0u;
    if (i < (size_t)n_act * kBinWords) {
        uint32_t f = (uint32_t)(i % 7);
        bins[i] = f == 0 ? 0u : (f <= 3 ? ~0u : 0u);
    }
}

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

template <class T>
hipError_t galloc(T** p, size_t n) {
    *p = nullptr;
    return hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
}
template <class T>
void gfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

}  // namespace

hipError_t gpu_build_bvh(const float* d_positions, uint32_t n, HostBvh& out, double* device_ms, hipStream_t stream) {
    out.nodes.clear();
    out.tri_index.clear();
    out.max_depth = 0;
    if (device_ms) *device_ms = 0.0;
    if (n == 0) return hipSuccess;
    const size_t max_nodes = 2 * (size_t)n;
    const size_t max_act = n / (kLeafTris + 1) + 1;
    float4 *cen = nullptr, *tmn = nullptr, *tmx = nullptr;
    int32_t *idx0 = nullptr, *idx1 = nullptr, *slot = nullptr, *target = nullptr, *ptab = nullptr;
    uint32_t *L = nullptr, *R = nullptr, *SL = nullptr, *SR = nullptr, *scan_tmp = nullptr, *scan_tot = nullptr;
    GNode* nodes = nullptr;
    unsigned long long* keys = nullptr;
    int32_t *act0 = nullptr, *act1 = nullptr, *split_axis = nullptr, *n_left = nullptr, *child_slot = nullptr;
    float* split_pos = nullptr;
    uint32_t *cb = nullptr, *bins = nullptr, *counters = nullptr;
    int32_t *internal = nullptr, *dfs_id = nullptr, *dfs_base = nullptr;
    BvhNode* d_out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipSuccess;
    std::vector<int32_t> level_begin;  // breadth-first id ranges of the levels
    int32_t n_nodes = 1;
    uint32_t host_counters[4] = {1u, 0u, 0u, 0u};
    int32_t* idx_final = nullptr;
    auto run = [&]() -> hipError_t {
#define PG_B(x)                         \
    do {                                \
        hipError_t _e = (x);            \
        if (_e != hipSuccess) return _e; \
    } while (0)
        PG_B(galloc(&cen, n)); PG_B(galloc(&tmn, n)); PG_B(galloc(&tmx, n));
        PG_B(galloc(&idx0, n)); PG_B(galloc(&idx1, n)); PG_B(galloc(&slot, n)); PG_B(galloc(&target, n));
        PG_B(galloc(&ptab, (size_t)n + 1));
        PG_B(galloc(&L, n)); PG_B(galloc(&R, n)); PG_B(galloc(&SL, n)); PG_B(galloc(&SR, n));
        PG_B(galloc(&scan_tmp, scan_tiles(n) + 1)); PG_B(galloc(&scan_tot, 1));
        PG_B(galloc(&nodes, max_nodes)); PG_B(galloc(&keys, 6 * max_nodes));
        PG_B(galloc(&act0, max_act)); PG_B(galloc(&act1, max_act));
        PG_B(galloc(&split_axis, max_act)); PG_B(galloc(&split_pos, max_act)); PG_B(galloc(&n_left, max_act));
        PG_B(galloc(&child_slot, max_nodes));
        PG_B(galloc(&cb, 6 * max_act)); PG_B(galloc(&bins, (size_t)kBinWords * max_act));
        PG_B(galloc(&counters, 4));
        PG_B(hipEventCreate(&e0)); PG_B(hipEventCreate(&e1));
        PG_B(hipEventRecord(e0, stream));
        hipLaunchKernelGGL(k_tri_prep, grid_for(n), dim3(kBlock), 0, stream, d_positions, n, cen, tmn, tmx, idx0,
                           idx1, slot);
        // root: all triangles, bounds over every triangle in input order (Geometry.h:436-441)
        GNode root{0, (int32_t)n, -1, 0, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        PG_B(hipMemcpyAsync(nodes, &root, sizeof(GNode), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_fill_u64, grid_for(6 * max_nodes), dim3(kBlock), 0, stream, keys, 6 * max_nodes, ~0ull);
        hipLaunchKernelGGL(k_fill_u32, grid_for(n), dim3(kBlock), 0, stream, (uint32_t*)target, (size_t)n, 0u);
        hipLaunchKernelGGL(k_bound_keys, grid_for(n), dim3(kBlock), 0, stream, target, idx0, tmn, tmx, n, keys);
        hipLaunchKernelGGL(k_bound_finalize, dim3(1), dim3(kBlock), 0, stream, keys, idx0, tmn, tmx, n, 0, 1, nodes);
        int32_t zero = 0;
        PG_B(hipMemcpyAsync(act0, &zero, 4, hipMemcpyHostToDevice, stream));
        PG_B(hipMemcpyAsync(counters, host_counters, sizeof(host_counters), hipMemcpyHostToDevice, stream));
        PG_B(hipGetLastError());
        level_begin.push_back(0);
        uint32_t n_act = n > (uint32_t)kLeafTris ? 1u : 0u;
        int32_t *in = idx0, *outb = idx1, *cur_act = act0, *next_act = act1;
        idx_final = idx0;
        int levels = 0;
        while (n_act > 0) {
            // a split takes at least one triangle off a node; traversal stacks hold 64
            // levels, so a build this deep is refused rather than run level by level
            if (++levels > kMaxLevels) return hipErrorInvalidValue;
            const size_t w = std::max<size_t>((size_t)n_act * kBinWords, 6 * (size_t)n_act);
            hipLaunchKernelGGL(k_level_init, grid_for(w), dim3(kBlock), 0, stream, cb, bins, n_act);
            hipLaunchKernelGGL(k_centroid_bounds, grid_for(n), dim3(kBlock), 0, stream, slot, in, cen, n, cb);
            hipLaunchKernelGGL(k_bins, grid_for(n), dim3(kBlock), 0, stream, slot, in, cen, tmn, tmx, n, cb, bins);
            hipLaunchKernelGGL(k_sah, grid_for(n_act), dim3(kBlock), 0, stream, cur_act, n_act, nodes, cb, bins,
                               split_axis, split_pos);
            hipLaunchKernelGGL(k_flags, grid_for(n), dim3(kBlock), 0, stream, slot, in, cen, n, split_axis, split_pos,
                               L, R);
            PG_B(exclusive_scan(L, SL, n, scan_tmp, scan_tot, stream));
            PG_B(exclusive_scan(R, SR, n, scan_tmp, scan_tot, stream));
            hipLaunchKernelGGL(k_split_counts, grid_for(n_act), dim3(kBlock), 0, stream, cur_act, n_act, nodes,
                               split_axis, L, SL, R, SR, n, n_left, ptab);
            hipLaunchKernelGGL(k_ptab, grid_for(n), dim3(kBlock), 0, stream, slot, n, cur_act, nodes, n_left, L, SR,
                               ptab);
            hipLaunchKernelGGL(k_partition, grid_for(n), dim3(kBlock), 0, stream, slot, n, cur_act, nodes, n_left, L,
                               SL, ptab, in, outb);
            uint32_t zero32 = 0;
            PG_B(hipMemcpyAsync(counters + 1, &zero32, 4, hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_children, grid_for(n_act), dim3(kBlock), 0, stream, cur_act, n_act, nodes, n_left,
                               counters, child_slot, next_act);
            hipLaunchKernelGGL(k_child_map, grid_for(n), dim3(kBlock), 0, stream, slot, n, cur_act, nodes, n_left,
                               child_slot, target, in, outb);
            hipLaunchKernelGGL(k_bound_keys, grid_for(n), dim3(kBlock), 0, stream, target, outb, tmn, tmx, n, keys);
            PG_B(hipGetLastError());
            PG_B(hipMemcpyAsync(host_counters, counters, sizeof(host_counters), hipMemcpyDeviceToHost, stream));
            PG_B(hipStreamSynchronize(stream));
            const int32_t begin = n_nodes, end = (int32_t)host_counters[0];
            if (end > begin) {
                hipLaunchKernelGGL(k_bound_finalize, grid_for((size_t)(end - begin)), dim3(kBlock), 0, stream, keys,
                                   outb, tmn, tmx, n, begin, end, nodes);
                level_begin.push_back(begin);
            }
            n_nodes = end;
            n_act = host_counters[1];
            idx_final = outb;
            std::swap(in, outb);
            std::swap(cur_act, next_act);
        }
        level_begin.push_back(n_nodes);
        // depth-first renumbering (levels bottom up for subtree sizes, top down for ids)
        PG_B(galloc(&internal, (size_t)n_nodes)); PG_B(galloc(&dfs_id, (size_t)n_nodes));
        PG_B(galloc(&dfs_base, (size_t)n_nodes)); PG_B(galloc(&d_out, (size_t)n_nodes));
        for (int lv = (int)level_begin.size() - 2; lv >= 0; --lv) {
            int32_t b = level_begin[lv], en = level_begin[lv + 1];
            hipLaunchKernelGGL(k_internal, grid_for((size_t)(en - b)), dim3(kBlock), 0, stream, nodes, b, en, internal);
        }
        int32_t root_ids[2] = {0, 1};
        PG_B(hipMemcpyAsync(dfs_id, &root_ids[0], 4, hipMemcpyHostToDevice, stream));
        PG_B(hipMemcpyAsync(dfs_base, &root_ids[1], 4, hipMemcpyHostToDevice, stream));
        for (size_t lv = 0; lv + 1 < level_begin.size(); ++lv) {
            int32_t b = level_begin[lv], en = level_begin[lv + 1];
            hipLaunchKernelGGL(k_dfs, grid_for((size_t)(en - b)), dim3(kBlock), 0, stream, nodes, b, en, internal,
                               dfs_id, dfs_base);
        }
        hipLaunchKernelGGL(k_write_nodes, grid_for((size_t)n_nodes), dim3(kBlock), 0, stream, nodes, n_nodes, dfs_id,
                           dfs_base, d_out);
        PG_B(hipGetLastError());
        PG_B(hipEventRecord(e1, stream));
        out.nodes.resize((size_t)n_nodes);
        out.tri_index.resize(n);
        PG_B(hipMemcpyAsync(out.nodes.data(), d_out, (size_t)n_nodes * sizeof(BvhNode), hipMemcpyDeviceToHost,
                            stream));
        PG_B(hipMemcpyAsync(out.tri_index.data(), idx_final, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
        PG_B(hipStreamSynchronize(stream));
        float ms = 0.f;
        PG_B(hipEventElapsedTime(&ms, e0, e1));
        if (device_ms) *device_ms = ms;
        out.max_depth = host_counters[2];
        return hipSuccess;
#undef PG_B
    };
    e = run();
    gfree(cen); gfree(tmn); gfree(tmx); gfree(idx0); gfree(idx1); gfree(slot); gfree(target); gfree(ptab);
    gfree(L); gfree(R); gfree(SL); gfree(SR); gfree(scan_tmp); gfree(scan_tot); gfree(nodes); gfree(keys);
    gfree(act0); gfree(act1); gfree(split_axis); gfree(split_pos); gfree(n_left); gfree(child_slot); gfree(cb);
    gfree(bins); gfree(counters); gfree(internal); gfree(dfs_id); gfree(dfs_base); gfree(d_out);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        out.nodes.clear();
        out.tri_index.clear();
    }
    return e;
}

}  // namespace pg
