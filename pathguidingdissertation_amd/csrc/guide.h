// guide.h — radiance cache of path-vertex records: point BVH over Morton-sorted
// records, exact k-nearest-neighbour queries and per-cell directional quadtrees.
//
// The reference keeps PathVertex records (Renderer.h:42-47) in a pointer-based
// median-split point BVH (Renderer.h:71-234) with no query API; the guiding design
// it outlines (Renderer.h:398-414) needs nearby-vertex lookups and a directional
// distribution. Here the records are sorted by 30-bit Morton code; leaves hold 8
// consecutive records (leaf <= 8 as Renderer.h:78) under an implicit balanced tree;
// a cell is a run of consecutive sorted records and owns one quadtree over the
// records' incident directions, refined where a quadrant holds more than 1% of the
// cell's flux, to depth 10.
#pragma once
#include <hip/hip_runtime.h>

#include "pg_common.h"

namespace pg {

constexpr int kLeafRecords = 8;
constexpr int kKnnMax = 16;
constexpr int kQtMaxDepth = 10;
constexpr unsigned long long kQtSplitDivisor = 100ull;  // refine above 1% of the cell's flux
constexpr double kQtFixedScale = 16777216.0;  // 2^24: flux sums in 64-bit fixed point

// 48-byte record in HBM: pos.xyz, normal.xyz, wi.xyz, Li.rgb.
struct alignas(16) Record {
    float v[12];
};
static_assert(sizeof(Record) == 48, "48-byte record");

struct alignas(16) QtNode {
    float e[4];       // flux of the four quadrants (qx + 2 qy)
    int32_t child[4]; // child node per quadrant, -1 for a leaf quadrant
};
static_assert(sizeof(QtNode) == 32, "32-byte quadtree node");

struct GuideView {
    uint32_t n_records;
    uint32_t n_leaf_pow2;       // P: leaves padded to a power of two
    uint32_t levels;            // log2(P)
    uint32_t cell_size;
    uint32_t n_cells;
    uint32_t qt_stride;         // nodes reserved per cell
    const float4* rec_pos;      // sorted: xyz, original index bits
    const float4* nodes;        // 2 * (2P - 1) float4: bmin, bmax
    const QtNode* qt;           // n_cells * qt_stride
    const float4* rec_dir;      // sorted: record direction xyz, flux (luminance of Li)
    const uint32_t* rec_key;    // sorted 30-bit Morton keys of the record positions
    const uint32_t* leaf_nbrs;  // per leaf: the kKnnMax records nearest its centroid
    float m_lo[3], m_inv;       // Morton frame of rec_key (scene bounding cube)
    float guide_prob;           // one-sample MIS probability of sampling the guide
};

// Cylindrical (equal-area) map between directions and [0,1)^2: u = phi / 2pi,
// v = (cos theta + 1) / 2 with cos theta = w.z; the Jacobian is 4 pi everywhere.
PG_HD void dir_to_uv(float x, float y, float z, float& u, float& v) {
    float phi = atan2f(y, x);
    if (phi < 0.f) phi += 6.283185307179586f;
    u = phi * 0.15915494309189535f;
    if (u >= 1.f) u = 0.f;
    v = (z + 1.f) * 0.5f;
    v = v < 0.f ? 0.f : (v >= 1.f ? 0.99999994f : v);
}
PG_HD void uv_to_dir(float u, float v, float& x, float& y, float& z) {
    float cz = 2.f * v - 1.f;
    float sz = sqrtf(fmaxf(0.f, 1.f - cz * cz));
    float phi = 6.283185307179586f * u;
    x = sz * cosf(phi);
    y = sz * sinf(phi);
    z = cz;
}

}  // namespace pg
