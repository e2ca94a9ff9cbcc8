// Shared internals of libfe_hip.so: the handle, error reporting and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "fe_hip.h"

struct fe_handle {
  int device;
  hipStream_t stream;
  int num_cus;
  // Two pinned host slots for polling solver status without blocking the queue.
  void* pinned;
};

int fe_set_error(int code, const char* fmt, ...);

// Coarse correction u += P Ainv P^T r on `stream` (fe_coarse.hip); rc, zc: n_agg * nc scratch values.
// state (may be null): the {done, need_ru} words of the solve's first segment, state_stride int32
// apart for the next segments (block-diagonal space); mode 0 runs while a segment iterates, 1 during
// its reliable update, 2 always.
int fe_coarse_correction(hipStream_t stream, const fe_coarse* C, int fp32, const void* Ainv,
                         const void* r, void* u, void* rc, void* zc, const int32_t* state,
                         int state_stride, int mode);

// Look up (once per device, with that device current) the device address of the file-scope scratch
// array of fe_coarse_apply_noalloc (fe_coarse.hip); fe_create calls it.
int fe_coarse_scratch_setup(int device);

#define FE_CHECK_HIP(call)                                                              \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fe_set_error((int)e_, "%s failed: %s", #call, hipGetErrorString(e_));      \
  } while (0)

#define FE_REQUIRE(cond, ...)                                                           \
  do {                                                                                  \
    if (!(cond)) return fe_set_error(FE_ERR_ARG, __VA_ARGS__);                          \
  } while (0)

// Check the launch of the last kernel on the handle's stream.
#define FE_CHECK_LAUNCH() FE_CHECK_HIP(hipGetLastError())

static inline int fe_set_device(fe_handle* h) {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != h->device) {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return fe_set_error((int)e, "hipSetDevice: %s", hipGetErrorString(e));
  }
  return FE_OK;
}

#define FE_ENTER(h)                                                                     \
  do {                                                                                  \
    if (!(h)) return fe_set_error(FE_ERR_ARG, "null handle");                           \
    int rc_ = fe_set_device(h);                                                         \
    if (rc_) return rc_;                                                                \
  } while (0)

__host__ __device__ static inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Upper-triangle row-major index of (i, j), i <= j, in an m x m symmetric matrix.
__host__ __device__ static inline int sym_index(int i, int j, int m) {
  return i * m - (i * (i - 1)) / 2 + (j - i);
}

// Remap a workgroup id so that the 8 XCDs of an MI355X each own a contiguous range of logical ids
// (workgroups are dispatched round-robin over the XCDs).
__device__ static inline int64_t xcd_remap(int64_t bid, int64_t nblocks) {
  const int64_t nx = 8;
  if (nblocks < nx || nblocks % nx) return bid;
  return (bid % nx) * (nblocks / nx) + bid / nx;
}
