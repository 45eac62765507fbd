// Handle, scratch memory and error reporting of the C ABI.
#include <cstdarg>
#include <cstring>

#include "common.hpp"

namespace ccs {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
  return status;
}

int Scratch::reserve(size_t need) {
  if (need <= bytes) return CCS_OK;
  if (ptr) {
    // The buffer may still be read by queued work of this handle.
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(CCS_ERR_HIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
    (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  size_t rounded = (need + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  hipError_t e = hipMalloc(&ptr, rounded);
  if (e != hipSuccess) {
    ptr = nullptr;
    return fail(CCS_ERR_ALLOC, "hipMalloc(%zu) failed: %s", rounded, hipGetErrorString(e));
  }
  bytes = rounded;
  return CCS_OK;
}

void Scratch::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

}  // namespace ccs

extern "C" const char* ccs_last_error(void) { return ccs::g_error; }

extern "C" const char* ccs_version(void) { return "ccs-mi355x 0.1 (gfx950)"; }

extern "C" int ccs_feature_stride(int C) { return ccs::feature_stride(C); }

extern "C" int ccs_create(ccs_handle_t* handle, void* hip_stream) {
  CCS_CHECK(handle != nullptr, CCS_ERR_ARG, "ccs_create: handle pointer is NULL");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return ccs::fail(CCS_ERR_HIP, "ccs_create: no HIP device: %s", hipGetErrorString(e));
  ccs_handle_s* h = new ccs_handle_s();
  h->stream = (hipStream_t)hip_stream;
  *handle = h;
  return CCS_OK;
}

extern "C" int ccs_destroy(ccs_handle_t h) {
  if (!h) return CCS_OK;
  (void)hipStreamSynchronize(h->stream);
  h->conv_w.release();
  h->crf_a.release();
  h->crf_msg.release();
  h->crf_small.release();
  h->crf_sums.release();
  h->misc.release();
  h->misc2.release();
  delete h;
  return CCS_OK;
}

extern "C" int ccs_set_stream(ccs_handle_t h, void* hip_stream) {
  CCS_CHECK(h != nullptr, CCS_ERR_ARG, "ccs_set_stream: NULL handle");
  h->stream = (hipStream_t)hip_stream;
  return CCS_OK;
}

extern "C" int ccs_synchronize(ccs_handle_t h) {
  CCS_CHECK(h != nullptr, CCS_ERR_ARG, "ccs_synchronize: NULL handle");
  CCS_HIP(hipStreamSynchronize(h->stream));
  return CCS_OK;
}
