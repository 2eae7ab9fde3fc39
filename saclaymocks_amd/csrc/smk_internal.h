// Host-side helpers shared by the C-ABI translation units: error state and
// HIP call checking.  Every exported entry point returns 0 on success or a
// negative SMK_ERR_* code and leaves a message for smk_last_error().
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/saclaymocks_hip.h"

namespace smk {
int set_error(int code, const char* fmt, ...);
}

#define SMK_HIP_CHECK(expr)                                                                 \
  do {                                                                                      \
    hipError_t smk_e_ = (expr);                                                             \
    if (smk_e_ != hipSuccess)                                                               \
      return smk::set_error(SMK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(smk_e_)); \
  } while (0)

#define SMK_REQUIRE(cond, code, ...)                         \
  do {                                                       \
    if (!(cond)) return smk::set_error((code), __VA_ARGS__); \
  } while (0)

#define SMK_LAUNCH_CHECK()                                                                     \
  do {                                                                                         \
    hipError_t smk_e_ = hipGetLastError();                                                     \
    if (smk_e_ != hipSuccess)                                                                  \
      return smk::set_error(SMK_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(smk_e_)); \
  } while (0)
