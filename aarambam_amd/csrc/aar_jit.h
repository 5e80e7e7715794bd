// Fused z kernels of the NG sum compiled at run time for one job's product schedule.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "aar_kernels.h"

struct JitZng {
  hipFunction_t fn = nullptr;
  int rows = 0, threads = 0;  // grid rows per block and threads per block
};

// rows per block and threads per block of the fused z kernel of M mode fields at H = N / 2
void jit_zng_shape(int fp64, int H, int M, int* rows, int* threads);
// source of the kernel for one schedule: groups (K + 1 offsets, in records) and recs as
// uploaded for k_z_ng (aar_job.cpp ng_factor)
std::string jit_zng_source(int fp64, int H, int M, int K, const int* groups, const NgRec4* recs);
// compile (once per device and source) and load; the compiler's log in *log
int jit_zng_kernel(int device, const std::string& src, JitZng* out, std::string* log);
// launch over the whole grid on stream st (same arguments as Pipeline::z_ng)
int jit_zng_launch(const JitZng& k, void* base, int M, int K, const Geom& g, const void* tw, hipStream_t st);
