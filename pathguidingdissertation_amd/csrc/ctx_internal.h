// ctx_internal.h — context accessors shared between the C-ABI translation units
// (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/pg_tracer.h"

namespace pg_internal {
hipStream_t stream(pg_ctx* c);
int cu_count(pg_ctx* c);
int device(pg_ctx* c);
pg_status fail(pg_ctx* c, pg_status code, const std::string& message);
}  // namespace pg_internal
