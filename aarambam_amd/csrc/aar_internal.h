// Library-internal entry points shared between the job and the serving code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/aarambam.h"

// aar_job_prepare with the job's device buffers carved from `ws` (ws_size bytes,
// at least aar_job_bytes) and its table uploads enqueued on `up`.
int aar_job_prepare_ws(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision, int schedule,
                       void* ws, size_t ws_size, hipStream_t up, aar_job** out);
// HIP device ordinal and stream of a context.
int aar_ctx_device(const aar_ctx* ctx);
hipStream_t aar_ctx_stream(const aar_ctx* ctx);
// hipMalloc with the library's status code (AAR_ERR_ALLOC, pointer nulled).
int aar_device_alloc(void** p, size_t bytes);
