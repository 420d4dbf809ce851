// ctx_internal.h — context accessors shared between the C-ABI translation units
// (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pg_tracer.h"

namespace pg_internal {
hipStream_t stream(pg_ctx* c);
int cu_count(pg_ctx* c);
int device(pg_ctx* c);
pg_status fail(pg_ctx* c, pg_status code, const std::string& message);

// Albedo textures of one uploaded scene in device memory (pg_scene_textures,
// pg_batch_scene_textures): RGBA32F texels, a (first texel, width, height) table and
// three (u, v) per triangle.
struct TextureSet {
    float4* texels = nullptr;
    int4* info = nullptr;
    float2* uv = nullptr;
    void release();
};
// Validates the arrays against the scene's triangle and material counts and uploads
// them (set is released first). Returns an empty string or the reason; *code is
// PG_ERR_INVALID for bad arguments, PG_ERR_NOMEM / PG_ERR_DEVICE for HIP failures.
std::string upload_textures(TextureSet& set, size_t n_tris, size_t n_mats, const float* uv,
                            const int32_t* material_texture, uint32_t n_materials, const float* texels,
                            const uint64_t* texel_offset, const int32_t* tex_width, const int32_t* tex_height,
                            uint32_t n_textures, pg_status* code);
}  // namespace pg_internal
