// Shared types and helpers for the MI355X NG-2LPT library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

// Status codes (AAR_OK, AAR_ERR_*) are part of the C-ABI header.
#include "../../include/aarambam.h"

#define AAR_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
              __LINE__);                                                           \
      return AAR_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#include "aar_device.h"
