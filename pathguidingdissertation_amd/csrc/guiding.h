// guiding.h — buffers and launchers of the record pass and the radiance cache.
#pragma once
#include <hip/hip_runtime.h>

#include "guide.h"

namespace pg {

// Per-path forward vertices of the record pass (Renderer.h:49-60), vertex-major:
// entry (depth, path) at depth * stride + path.
struct RecordState {
    float4* vpos;  // hit point, store flag in .w
    float4* vnrm;  // shading normal
    float4* vwi;   // sampled incident direction, sampling pdf q in .w
    float4* vwo;   // outgoing direction (-ray), BSDF pdf of wi in .w
    float4* vw;    // bsdf weight f |cos| / (pdf rrp)
    float4* ve;    // emission (Le or background)
    float4* vme;   // MIS-weighted emission (Renderer.h:433-457, 503-525)
    float4* vd;    // next-event estimate at the vertex (Renderer.h:463)
    uint32_t* nv;  // vertices per path
    uint32_t stride;
    int32_t li_nee;  // 1: records carry Li with NEE (the pixelColour recursion)
};

struct CacheBuffers {
    uint32_t* keys;
    uint32_t* vals;
    uint32_t* keys_tmp;
    uint32_t* vals_tmp;
    uint32_t* hist;
    uint32_t* scan_tmp;
    float4* pos;       // sorted record positions + original index
    float4* dir;       // sorted record directions + flux (luminance of Li)
    float4* nodes;     // 2 float4 per node, 2P-1 nodes
    QtNode* qt;
    unsigned long long* qt_acc;
    uint32_t leaf_pow2;
    uint32_t levels;
};

uint32_t scan_tiles(uint32_t n);
uint32_t sort_tiles(uint32_t n);
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tile_sums, uint32_t* total,
                          hipStream_t stream);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                            int bits, uint32_t* hist, uint32_t* scan_tmp, hipStream_t stream);
hipError_t launch_backprop_count(const RecordState& rs, uint32_t n_paths, uint32_t* counts, hipStream_t stream);
struct TrainAux;
hipError_t launch_backprop_write(const RecordState& rs, uint32_t n_paths, const uint32_t* offsets, uint32_t base,
                                 Record* out, TrainAux* aux, uint32_t capacity, float4* radiance, hipStream_t stream);
hipError_t launch_cache_build(const Record* rec, uint32_t n, float3 lo, float3 inv_ext, CacheBuffers& cb,
                              uint32_t cell_size, uint32_t qt_stride, hipStream_t stream);
hipError_t launch_knn(const GuideView& g, const float* q, uint32_t nq, uint32_t k, uint32_t* idx, float* d2,
                      hipStream_t stream);
hipError_t launch_guide_eval(const GuideView& g, const float* pos, const float* u3, const float* wi_eval, uint32_t n,
                             float* wi_out, float* pdf_sample, float* pdf_eval, hipStream_t stream);

}  // namespace pg
