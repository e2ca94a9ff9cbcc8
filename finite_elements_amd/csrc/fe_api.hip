// Context, error reporting and small utility entry points of libfe_hip.so.
#include "fe_internal.h"

#include <new>

static thread_local std::string g_last_error;

int fe_set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

extern "C" int fe_version(void) { return 1; }

extern "C" const char* fe_last_error(void) { return g_last_error.c_str(); }

extern "C" int fe_create(fe_handle** out, int device, void* stream) {
  if (!out) return fe_set_error(FE_ERR_ARG, "fe_create: null output pointer");
  *out = nullptr;
  int ndev = 0;
  FE_CHECK_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fe_set_error(FE_ERR_ARG, "fe_create: device %d out of range (%d devices)", device, ndev);
  FE_CHECK_HIP(hipSetDevice(device));
  fe_handle* h = new (std::nothrow) fe_handle();
  if (!h) return fe_set_error(FE_ERR_NOMEM, "fe_create: out of host memory");
  h->device = device;
  h->stream = reinterpret_cast<hipStream_t>(stream);
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete h;
    return fe_set_error((int)e, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  }
  h->num_cus = prop.multiProcessorCount;
  e = hipHostMalloc(&h->pinned, 4096, hipHostMallocDefault);
  if (e != hipSuccess) {
    delete h;
    return fe_set_error((int)e, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  // device address of the allocation-free entry point's scratch array on this device
  int rc = fe_coarse_scratch_setup(device);
  if (rc) {
    hipHostFree(h->pinned);
    delete h;
    return rc;
  }
  *out = h;
  return FE_OK;
}

extern "C" int fe_destroy(fe_handle* h) {
  if (!h) return FE_OK;
  fe_set_device(h);
  hipStreamSynchronize(h->stream);
  if (h->pinned) hipHostFree(h->pinned);
  delete h;
  return FE_OK;
}

extern "C" int fe_synchronize(fe_handle* h) {
  FE_ENTER(h);
  FE_CHECK_HIP(hipStreamSynchronize(h->stream));
  return FE_OK;
}

__global__ void k_cast_f64_f32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = (float)src[i];
}

extern "C" int fe_cast_f64_to_f32(fe_handle* h, int64_t n, const double* src, float* dst) {
  FE_ENTER(h);
  FE_REQUIRE(n >= 0 && (n == 0 || (src && dst)), "fe_cast_f64_to_f32: bad arguments");
  if (n == 0) return FE_OK;
  int64_t grid = ceil_div(n, 256);
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_cast_f64_f32, dim3((unsigned)grid), dim3(256), 0, h->stream, n, src, dst);
  FE_CHECK_LAUNCH();
  return FE_OK;
}
