// bvh_build.h — the binned-SAH BVH build on the GPU (bvh_build.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "host_prep.hpp"

namespace pg {
// Builds the BVH of n triangles (9 floats each, device memory) in the reference's
// topology, node numbering and triangle order (BVHNode::build, Geometry.h:274-446),
// the same arrays build_bvh produces on the host, and copies them into `out`.
// device_ms: GPU time of the build (events on `stream`). Allocates and frees its
// own device buffers (~1 KB per triangle at most).
hipError_t gpu_build_bvh(const float* d_positions, uint32_t n, HostBvh& out, double* device_ms, hipStream_t stream);
}  // namespace pg
