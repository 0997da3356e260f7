// noalloc.hip -- md_probe_fixed_point_noalloc: the fixed-point probe for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).
//
// md_probe_fixed_point stages its input and both outputs in three device
// buffers it allocates and frees on every call.  This variant stages them in
// `scratch`, a file-scope device array of fixed size that each GPU the
// library runs on holds until the process ends; it calls no hipMalloc or
// hipFree and no allocating library routine.  In host code `scratch` names a
// host-side stand-in, so the array's device address is taken once per device
// with hipGetSymbolAddress (with that device current) and kept; kernels and
// copies use that address only.  Every call on one GPU shares the array:
// calls that overlap on one GPU race.  Nothing in it needs to start at zero:
// a call writes every staged value it reads back.
#include <string>

#include "engine.hpp"

#define MD_NOALLOC_MAX_VALUES 65536  // values per call
#define MD_NOALLOC_MAX_DEVICES 64

// staging: f32 inputs, Q32.32 outputs, f32 read-backs (1 MiB)
static __device__ __align__(256) unsigned char
    scratch[MD_NOALLOC_MAX_VALUES * (sizeof(float) + sizeof(int64_t) + sizeof(float))];

// each device's address of `scratch`, set on the first call on that device
static unsigned char *g_scratch_ptr[MD_NOALLOC_MAX_DEVICES] = {};

static unsigned char *scratch_on_current_device() {
  int dev = 0;
  MD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MD_NOALLOC_MAX_DEVICES)
    throw md_error(MD_ERR_INVALID_ARGUMENT, "device ordinal beyond the library's table of staging arrays");
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw md_error(MD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  return g_scratch_ptr[dev];
}

__global__ void k_probe_fixed_staged(const float *in, int n, int64_t *fixed, float *back) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t f = md_to_fixed(in[i]);
  fixed[i] = f;
  back[i] = md_from_fixed(f);
}

extern "C" md_err md_probe_fixed_point_noalloc(const float *values, int32_t n, int64_t *fixed, float *back) {
  try {
    if (n < 0 || (n > 0 && (!values || !fixed || !back)))
      throw md_error(MD_ERR_INVALID_ARGUMENT, "bad probe arguments");
    if (n > MD_NOALLOC_MAX_VALUES)
      throw md_error(MD_ERR_INVALID_ARGUMENT, "md_probe_fixed_point_noalloc takes at most " +
                                                  std::to_string(MD_NOALLOC_MAX_VALUES) + " values per call, got " +
                                                  std::to_string(n));
    if (n == 0) return MD_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw md_error(MD_ERR_NO_DEVICE, "no HIP device");
    unsigned char *base = scratch_on_current_device();
    float *din = reinterpret_cast<float *>(base);
    int64_t *dfix = reinterpret_cast<int64_t *>(base + sizeof(float) * MD_NOALLOC_MAX_VALUES);
    float *dback = reinterpret_cast<float *>(base + (sizeof(float) + sizeof(int64_t)) * MD_NOALLOC_MAX_VALUES);
    MD_HIP(hipMemcpy(din, values, sizeof(float) * n, hipMemcpyHostToDevice));
    k_probe_fixed_staged<<<(n + 255) / 256, 256>>>(din, n, dfix, dback);
    MD_HIP(hipGetLastError());
    MD_HIP(hipMemcpy(fixed, dfix, sizeof(int64_t) * n, hipMemcpyDeviceToHost));
    MD_HIP(hipMemcpy(back, dback, sizeof(float) * n, hipMemcpyDeviceToHost));
  } catch (const md_error &e) {
    md_set_last_error(e.message);
    return e.code;
  }
  md_set_last_error(std::string());
  return MD_OK;
}
