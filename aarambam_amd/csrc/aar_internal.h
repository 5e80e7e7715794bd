// Library-internal entry points shared between the job and the serving code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/aarambam.h"

// aar_job_prepare with the job's own buffers carved from `ws` (ws_size bytes, at least
// aar_job_ws_bytes) and its uploads enqueued on `up`.  plan_in_ws = 0: the tables come from the
// context's plan cache (device-resident across jobs); 1: they are laid out at the start of `ws`
// too, so nothing is allocated.
// out_pos / out_vel (both or neither): device arrays of 3 npart floats that receive the particles
// instead of buffers of the job's own.
int aar_job_prepare_ws(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision, int schedule,
                       void* ws, size_t ws_size, hipStream_t up, int plan_in_ws, aar_job** out, float* out_pos,
                       float* out_vel);
// Workspace bytes aar_job_prepare_ws carves for a job (-1 if the job is invalid).
long long aar_job_ws_bytes(const aar_params* p, const aar_tables* t, int precision, int schedule, int plan_in_ws);
// HIP device ordinal and stream of a context.
int aar_ctx_device(const aar_ctx* ctx);
hipStream_t aar_ctx_stream(const aar_ctx* ctx);
// hipMalloc with the library's status code (AAR_ERR_ALLOC, pointer nulled).
int aar_device_alloc(void** p, size_t bytes);
