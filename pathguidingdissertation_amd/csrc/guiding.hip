// guiding.hip — record pass back-propagation, device-wide scan and radix sort, the
// point-BVH radiance cache, exact kNN and per-cell quadtree construction (gfx950).
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "guide_device.h"
#include "guiding.h"
#include "mlp.h"

namespace pg {

// ------------------------------------------------------------------ exclusive scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    // Hillis-Steele over 256 lanes in LDS
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        uint32_t x = t >= off ? lds[t - off] : 0u;
        __syncthreads();
        lds[t] += x;
        __syncthreads();
    }
    total = lds[kScanBlock - 1];
    uint32_t incl = lds[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n,
                                                            uint32_t* tile_sums) {
    __shared__ uint32_t lds[kScanBlock];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t pre = block_exclusive_scan(sum, lds, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Scans the tile sums in one block (sequential chunks of 256) and writes the total.
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* sums, uint32_t n, uint32_t* total_out) {
    __shared__ uint32_t lds[kScanBlock];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < n; b += kScanBlock) {
        uint32_t i = b + threadIdx.x;
        uint32_t v = i < n ? sums[i] : 0u;
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(v, lds, tot);
        if (i < n) sums[i] = pre + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* tile_sums) {
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t add = tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] += add;
}

uint32_t scan_tiles(uint32_t n) { return (n + kScanTile - 1) / kScanTile; }

hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tile_sums, uint32_t* total,
                          hipStream_t stream) {
    if (n == 0) return hipSuccess;
    uint32_t nt = scan_tiles(n);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanBlock), 0, stream, in, out, n, tile_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, stream, tile_sums, nt, total);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kScanBlock), 0, stream, out, n, tile_sums);
    return hipGetLastError();
}

// ------------------------------------------------------------------ LSD radix sort
// 8-bit digits; per pass: tile histograms -> scan over (digit, tile) -> stable
// scatter. Inside a tile, keys are ranked in 256-key chunks in order: the lanes of
// a wavefront with the same digit are found with eight ballots, waves are ordered
// through LDS counters, and a running per-digit offset carries across chunks.
constexpr int kSortBlock = 256;
constexpr int kSortChunks = 16;
constexpr int kSortTile = kSortBlock * kSortChunks;

__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const uint32_t* keys, uint32_t n, int shift,
                                                           uint32_t* hist, uint32_t n_tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    for (int c = 0; c < kSortChunks; ++c) {
        uint32_t i = base + c * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(const uint32_t* keys, const uint32_t* vals,
                                                              uint32_t* okeys, uint32_t* ovals, uint32_t n, int shift,
                                                              const uint32_t* offsets, uint32_t n_tiles) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t cnt[kSortBlock / 64][256];
    const int t = threadIdx.x, wave = t >> 6;
    running[t] = offsets[t * n_tiles + blockIdx.x];
    for (int w = 0; w < kSortBlock / 64; ++w) cnt[w][t] = 0;
    __syncthreads();
    uint32_t base = blockIdx.x * kSortTile;
    const int lane = __lane_id();
    const unsigned long long below = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    for (int c = 0; c < kSortChunks; ++c) {
        uint32_t i = base + c * kSortBlock + t;
        bool valid = i < n;
        uint32_t key = valid ? keys[i] : 0u;
        uint32_t val = valid ? vals[i] : 0u;
        uint32_t digit = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (digit >> b) & 1u;
            unsigned long long m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        if (!valid) peers = 0ull;
        uint32_t rank = (uint32_t)__popcll(peers & below);
        int leader = valid ? (__ffsll((long long)peers) - 1) : -1;
        if (valid && lane == leader) cnt[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[digit] + rank;
            for (int w = 0; w < wave; ++w) pos += cnt[w][digit];
            okeys[pos] = key;
            ovals[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kSortBlock / 64; ++w) { add += cnt[w][t]; cnt[w][t] = 0; }
        running[t] += add;
        __syncthreads();
    }
}

uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                            int bits, uint32_t* hist, uint32_t* scan_tmp, hipStream_t stream) {
    if (n <= 1) return hipSuccess;
    uint32_t nt = sort_tiles(n);
    int passes = (bits + 7) / 8;
    uint32_t *ka = keys, *va = vals, *kb = keys_tmp, *vb = vals_tmp;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_sort_hist, dim3(nt), dim3(kSortBlock), 0, stream, ka, n, shift, hist, nt);
        hipError_t e = exclusive_scan(hist, hist, 256 * nt, scan_tmp, nullptr, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sort_scatter, dim3(nt), dim3(kSortBlock), 0, stream, ka, va, kb, vb, n, shift, hist, nt);
        uint32_t* t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    if (ka != keys) {
        (void)hipMemcpyAsync(keys, ka, (size_t)n * 4, hipMemcpyDeviceToDevice, stream);
        (void)hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, stream);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ record pass back-propagation
// RayTracer::guidedPath backward loop (Renderer.h:379-392): Li_i = G_{i+1},
// G_i = emission_i + bsdfWeight_i * G_{i+1}; a vertex is kept when it sampled a
// non-specular continuation and Li is finite with positive luminance.
__device__ __forceinline__ C3 backprop_step(const RecordState& rs, size_t o, C3 G) {
    float4 w = rs.vw[o];
    if (rs.li_nee) {
        // pixelColour = misEmission + directLighting + bsdfWeight * pixelColour (Renderer.h:391)
        float4 me = rs.vme[o], d = rs.vd[o];
        return (c3(me.x, me.y, me.z) + c3(d.x, d.y, d.z)) + (c3(w.x, w.y, w.z) * G);
    }
    float4 e = rs.ve[o];
    return c3(e.x, e.y, e.z) + (c3(w.x, w.y, w.z) * G);
}

__global__ void k_backprop_count(RecordState rs, uint32_t n_paths, uint32_t* counts) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    uint32_t nv = rs.nv[p];
    C3 G = c3(0.f, 0.f, 0.f);
    uint32_t c = 0;
    for (int i = (int)nv - 1; i >= 0; --i) {
        size_t o = (size_t)i * rs.stride + p;
        float4 vp = rs.vpos[o];
        if (vp.w != 0.f && finite3(G) && lum(G) > 0.f) ++c;
        G = backprop_step(rs, o, G);
    }
    counts[p] = c;
}

// Records of one pass in (path, depth) order; `aux` (nullable) receives each record's
// training data: (wo, lum(Li * f|cos|/(q rr))) and (BSDF pdf of wi, q).
__global__ void k_backprop_write(RecordState rs, uint32_t n_paths, const uint32_t* offsets, uint32_t base,
                                 Record* out, TrainAux* aux, uint32_t capacity, float4* radiance) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    uint32_t nv = rs.nv[p];
    C3 G = c3(0.f, 0.f, 0.f);
    uint32_t w_at = base + offsets[p];
    for (int i = (int)nv - 1; i >= 0; --i) {
        size_t o = (size_t)i * rs.stride + p;
        float4 vp = rs.vpos[o];
        if (vp.w != 0.f && finite3(G) && lum(G) > 0.f) {
            if (w_at < capacity) {
                float4 vn = rs.vnrm[o], vw = rs.vwi[o];
                float4* dst = reinterpret_cast<float4*>(out + w_at);
                dst[0] = make_float4(vp.x, vp.y, vp.z, vn.x);
                dst[1] = make_float4(vn.y, vn.z, vw.x, vw.y);
                dst[2] = make_float4(vw.z, G.r, G.g, G.b);
                if (aux) {
                    float4 wo = rs.vwo[o], bw = rs.vw[o];
                    float w = lum(G * c3(bw.x, bw.y, bw.z));
                    aux[w_at].a = make_float4(wo.x, wo.y, wo.z, w);
                    aux[w_at].b = make_float4(wo.w, vw.w, 0.f, 0.f);
                }
            }
            ++w_at;
        }
        G = backprop_step(rs, o, G);
    }
    // the record pass returns G_0 as the sample value (Renderer.h:393); with NEE in the
    // records it returns the NEE + MIS estimate instead (Renderer.h:394)
    radiance[p] = make_float4(G.r, G.g, G.b, 0.f);
}

hipError_t launch_backprop_count(const RecordState& rs, uint32_t n_paths, uint32_t* counts, hipStream_t stream) {
    hipLaunchKernelGGL(k_backprop_count, dim3((n_paths + 255) / 256), dim3(256), 0, stream, rs, n_paths, counts);
    return hipGetLastError();
}
hipError_t launch_backprop_write(const RecordState& rs, uint32_t n_paths, const uint32_t* offsets, uint32_t base,
                                 Record* out, TrainAux* aux, uint32_t capacity, float4* radiance, hipStream_t stream) {
    hipLaunchKernelGGL(k_backprop_write, dim3((n_paths + 255) / 256), dim3(256), 0, stream, rs, n_paths, offsets, base,
                       out, aux, capacity, radiance);
    return hipGetLastError();
}

// ------------------------------------------------------------------ cache build
__device__ __forceinline__ uint32_t expand10(uint32_t x) {
    x &= 1023u;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void k_morton(const Record* rec, uint32_t n, float3 lo, float3 inv_ext, uint32_t* keys, uint32_t* vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* v = rec[i].v;
    int qx = min(max((int)((v[0] - lo.x) * inv_ext.x * 1024.f), 0), 1023);
    int qy = min(max((int)((v[1] - lo.y) * inv_ext.y * 1024.f), 0), 1023);
    int qz = min(max((int)((v[2] - lo.z) * inv_ext.z * 1024.f), 0), 1023);
    keys[i] = (expand10((uint32_t)qx) << 2) | (expand10((uint32_t)qy) << 1) | expand10((uint32_t)qz);
    vals[i] = i;
}

// sorted position s <- record order[s]: position + original index, direction + flux
__global__ void k_gather_sorted(const Record* rec, const uint32_t* order, uint32_t n, float4* pos, float4* dir) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t i = order[s];
    const float* v = rec[i].v;
    pos[s] = make_float4(v[0], v[1], v[2], __uint_as_float(i));
    float l = ((0.2126f * v[9]) + (0.7152f * v[10])) + (0.0722f * v[11]);
    dir[s] = make_float4(v[6], v[7], v[8], l);
}

__global__ void k_leaf_bounds(const float4* pos, uint32_t n, uint32_t P, float4* nodes) {
    uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= P) return;
    float inf = __int_as_float(0x7f800000);
    float4 lo = make_float4(inf, inf, inf, 0.f), hi = make_float4(-inf, -inf, -inf, 0.f);
    uint32_t b = leaf * kLeafRecords, e = min(b + (uint32_t)kLeafRecords, n);
    for (uint32_t s = b; s < e; ++s) {
        float4 p = pos[s];
        lo.x = fminf(lo.x, p.x); lo.y = fminf(lo.y, p.y); lo.z = fminf(lo.z, p.z);
        hi.x = fmaxf(hi.x, p.x); hi.y = fmaxf(hi.y, p.y); hi.z = fmaxf(hi.z, p.z);
    }
    uint32_t node = P - 1 + leaf;
    nodes[2 * node] = lo;
    nodes[2 * node + 1] = hi;
}

__global__ void k_level_bounds(float4* nodes, uint32_t first, uint32_t count) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    uint32_t node = first + k;
    uint32_t c1 = 2 * node + 1, c2 = c1 + 1;
    float4 a0 = nodes[2 * c1], a1 = nodes[2 * c1 + 1], b0 = nodes[2 * c2], b1 = nodes[2 * c2 + 1];
    nodes[2 * node] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
    nodes[2 * node + 1] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
}

// One workgroup per cell. Flux sums are 64-bit fixed point so that the tree does not
// depend on the order of the atomics.
__global__ __launch_bounds__(256) void k_qt_build(const float4* dir, uint32_t n, uint32_t cell_size, QtNode* qt,
                                                  uint32_t stride, unsigned long long* acc) {
    __shared__ unsigned long long total;
    __shared__ int n_nodes, lvl_begin, lvl_end;
    const uint32_t cell = blockIdx.x;
    const uint32_t b = cell * cell_size, e = min(b + cell_size, n);
    QtNode* nodes = qt + (size_t)cell * stride;
    unsigned long long* ac = acc + (size_t)cell * stride * 4;  // per node, per quadrant
    if (threadIdx.x == 0) {
        total = 0ull;
        n_nodes = 1;
        lvl_begin = 0;
        lvl_end = 1;
    }
    for (uint32_t k = threadIdx.x; k < 4; k += blockDim.x) ac[k] = 0ull;
    __syncthreads();
    // root level
    for (uint32_t s = b + threadIdx.x; s < e; s += blockDim.x) {
        float4 d = dir[s];
        if (!(d.w > 0.f) || !isfinite(d.w)) continue;
        unsigned long long f = (unsigned long long)((double)d.w * kQtFixedScale + 0.5);
        float u, v;
        dir_to_uv(d.x, d.y, d.z, u, v);
        int q = (u >= 0.5f ? 1 : 0) + 2 * (v >= 0.5f ? 1 : 0);
        atomicAdd(&ac[q], f);
        atomicAdd(&total, f);
    }
    __syncthreads();
    for (int depth = 0; depth < kQtMaxDepth; ++depth) {
        // finalize energies of this level's nodes and create children
        if (threadIdx.x == 0) {
            // a quadrant is refined when it holds more than 1% of the cell's flux:
            // f > total / 100 exactly (integer flux, so f > floor(total / 100))
            unsigned long long thr = total / kQtSplitDivisor;
            int begin = lvl_begin, end = lvl_end, nn = n_nodes;
            for (int nd = begin; nd < end; ++nd) {
                for (int q = 0; q < 4; ++q) {
                    unsigned long long f = ac[nd * 4 + q];
                    nodes[nd].e[q] = (float)((double)f / kQtFixedScale);
                    int ch = -1;
                    if (depth + 1 < kQtMaxDepth && total > 0ull && f > thr && nn < (int)stride) {
                        ch = nn++;
                        ac[ch * 4 + 0] = ac[ch * 4 + 1] = ac[ch * 4 + 2] = ac[ch * 4 + 3] = 0ull;
                        for (int k = 0; k < 4; ++k) {
                            nodes[ch].e[k] = 0.f;
                            nodes[ch].child[k] = -1;
                        }
                    }
                    nodes[nd].child[q] = ch;
                }
            }
            lvl_begin = end;
            lvl_end = nn;
            n_nodes = nn;
        }
        __syncthreads();
        if (lvl_begin == lvl_end) break;
        // accumulate record flux into the new nodes' quadrants
        for (uint32_t s = b + threadIdx.x; s < e; s += blockDim.x) {
            float4 d = dir[s];
            if (!(d.w > 0.f) || !isfinite(d.w)) continue;
            float u, v;
            dir_to_uv(d.x, d.y, d.z, u, v);
            int node = 0;
            float ox = 0.f, oy = 0.f, size = 1.f;
            int nd_depth = 0;
            while (true) {
                float half = size * 0.5f;
                int qx = (u >= ox + half) ? 1 : 0, qy = (v >= oy + half) ? 1 : 0;
                int ch = nodes[node].child[qx + 2 * qy];
                if (ch < 0) break;
                node = ch;
                ox += qx * half;
                oy += qy * half;
                size = half;
                ++nd_depth;
            }
            if (node >= lvl_begin && node < lvl_end) {
                float half = size * 0.5f;
                int qx = (u >= ox + half) ? 1 : 0, qy = (v >= oy + half) ? 1 : 0;
                unsigned long long f = (unsigned long long)((double)d.w * kQtFixedScale + 0.5);
                atomicAdd(&ac[node * 4 + qx + 2 * qy], f);
            }
        }
        __syncthreads();
    }
    // mark unused nodes of the pool
    if (threadIdx.x == 0) {
        for (int nd = lvl_begin; nd < lvl_end; ++nd) {
            for (int q = 0; q < 4; ++q) {
                nodes[nd].e[q] = (float)((double)ac[nd * 4 + q] / kQtFixedScale);
                nodes[nd].child[q] = -1;
            }
        }
    }
}

hipError_t launch_cache_build(const Record* rec, uint32_t n, float3 lo, float3 inv_ext, CacheBuffers& cb,
                              uint32_t cell_size, uint32_t qt_stride, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, stream, rec, n, lo, inv_ext, cb.keys, cb.vals);
    hipError_t e = radix_sort_pairs(cb.keys, cb.vals, cb.keys_tmp, cb.vals_tmp, n, 30, cb.hist, cb.scan_tmp, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather_sorted, dim3((n + 255) / 256), dim3(256), 0, stream, rec, cb.vals, n, cb.pos, cb.dir);
    uint32_t P = cb.leaf_pow2;
    hipLaunchKernelGGL(k_leaf_bounds, dim3((P + 255) / 256), dim3(256), 0, stream, cb.pos, n, P, cb.nodes);
    for (int level = (int)cb.levels - 1; level >= 0; --level) {
        uint32_t first = (1u << level) - 1u, count = 1u << level;
        hipLaunchKernelGGL(k_level_bounds, dim3((count + 255) / 256), dim3(256), 0, stream, cb.nodes, first, count);
    }
    uint32_t n_cells = (n + cell_size - 1) / cell_size;
    hipLaunchKernelGGL(k_qt_build, dim3(n_cells), dim3(256), 0, stream, cb.dir, n, cell_size, cb.qt, qt_stride,
                       cb.qt_acc);
    return hipGetLastError();
}

// ------------------------------------------------------------------ query kernels (API / tests)
constexpr int kKnnBlock = 128;
constexpr int kKnnStack = 32;

__global__ __launch_bounds__(kKnnBlock) void k_knn(GuideView g, const float* q, uint32_t nq, uint32_t k,
                                                   uint32_t* idx, float* d2) {
    __shared__ int sn[kKnnStack * kKnnBlock];
    __shared__ float sd[kKnnStack * kKnnBlock];
    uint32_t i = blockIdx.x * kKnnBlock + threadIdx.x;
    if (i >= nq) return;
    float bd[kKnnMax];
    uint32_t bs[kKnnMax];
    F3 qq{q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    knn_query<kKnnMax>(g, qq, bd, bs, sn + threadIdx.x, sd + threadIdx.x, kKnnBlock);
#pragma unroll
    for (int j = 0; j < kKnnMax; ++j) {
        if ((uint32_t)j < k) {
            idx[(size_t)i * k + j] = bs[j] == 0xffffffffu ? 0xffffffffu : orig_index(g, bs[j]);
            d2[(size_t)i * k + j] = bs[j] == 0xffffffffu ? 3.402823466e38f : bd[j];
        }
    }
}

hipError_t launch_knn(const GuideView& g, const float* q, uint32_t nq, uint32_t k, uint32_t* idx, float* d2,
                      hipStream_t stream) {
    if (nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_knn, dim3((nq + kKnnBlock - 1) / kKnnBlock), dim3(kKnnBlock), 0, stream, g, q, nq, k, idx, d2);
    return hipGetLastError();
}

// Guide distribution at a point: the mixture of the quadtrees of the cells its k
// nearest records belong to (each neighbour weighted 1/k).
__global__ __launch_bounds__(kKnnBlock) void k_guide_eval(GuideView g, const float* pos, const float* u3,
                                                          const float* wi_eval, uint32_t n, float* wi_out,
                                                          float* pdf_sample, float* pdf_eval) {
    __shared__ int sn[kKnnStack * kKnnBlock];
    __shared__ float sd[kKnnStack * kKnnBlock];
    uint32_t i = blockIdx.x * kKnnBlock + threadIdx.x;
    if (i >= n) return;
    float bd[kKnnMax];
    uint32_t bs[kKnnMax];
    F3 x{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    knn_query<kKnnMax>(g, x, bd, bs, sn + threadIdx.x, sd + threadIdx.x, kKnnBlock);
    int kk = 0;
    while (kk < kKnnMax && bs[kk] != 0xffffffffu) ++kk;
    if (kk == 0) {
        wi_out[3 * i] = wi_out[3 * i + 1] = 0.f;
        wi_out[3 * i + 2] = 1.f;
        pdf_sample[i] = 0.f;
        pdf_eval[i] = 0.f;
        return;
    }
    int j = min((int)(u3[3 * i] * (float)kk), kk - 1);
    const QtNode* tree = g.qt + (size_t)(bs[j] / g.cell_size) * g.qt_stride;
    float u, v, pu;
    qt_sample_uv(tree, u3[3 * i + 1], u3[3 * i + 2], u, v, pu);
    float dx, dy, dz;
    uv_to_dir(u, v, dx, dy, dz);
    wi_out[3 * i] = dx; wi_out[3 * i + 1] = dy; wi_out[3 * i + 2] = dz;
    float uu, vv, ue, ve;
    dir_to_uv(dx, dy, dz, uu, vv);
    dir_to_uv(wi_eval[3 * i], wi_eval[3 * i + 1], wi_eval[3 * i + 2], ue, ve);
    float ps = 0.f, pe = 0.f;
    for (int m = 0; m < kk; ++m) {
        const QtNode* t = g.qt + (size_t)(bs[m] / g.cell_size) * g.qt_stride;
        ps += qt_pdf_uv(t, uu, vv);
        pe += qt_pdf_uv(t, ue, ve);
    }
    const float inv4pi = 0.07957747154594767f;
    pdf_sample[i] = ps / (float)kk * inv4pi;
    pdf_eval[i] = pe / (float)kk * inv4pi;
}

hipError_t launch_guide_eval(const GuideView& g, const float* pos, const float* u3, const float* wi_eval, uint32_t n,
                             float* wi_out, float* pdf_sample, float* pdf_eval, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_eval, dim3((n + kKnnBlock - 1) / kKnnBlock), dim3(kKnnBlock), 0, stream, g, pos, u3,
                       wi_eval, n, wi_out, pdf_sample, pdf_eval);
    return hipGetLastError();
}

}  // namespace pg
