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

int Scratch::reserve(size_t need, hipStream_t st) {
  if (need <= bytes) return CCS_OK;
  // Growing is stream-ordered (the old buffer is released after the work
  // queued on st that may still read it), so it blocks no other stream; it
  // cannot be recorded into a graph, which the caller learns here rather than
  // from a failed capture.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cs);
  if (e != hipSuccess) return fail(CCS_ERR_HIP, "hipStreamIsCapturing: %s", hipGetErrorString(e));
  if (cs != hipStreamCaptureStatusNone)
    return fail(CCS_ERR_ARG,
                "the handle's scratch must grow to %zu bytes while its stream is being captured: "
                "call ccs_reserve (or the same entry point once, uncaptured) at this size before capture",
                need);
  if (ptr) {
    e = hipFreeAsync(ptr, st);
    if (e != hipSuccess) return fail(CCS_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
    ptr = nullptr;
    bytes = 0;
  }
  size_t rounded = (need + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
  e = hipMallocAsync(&ptr, rounded, st);
  if (e != hipSuccess) {
    ptr = nullptr;
    return fail(CCS_ERR_ALLOC, "hipMallocAsync(%zu) failed: %s", rounded, hipGetErrorString(e));
  }
  bytes = rounded;
  return CCS_OK;
}

void Scratch::release(hipStream_t st) {
  if (ptr) (void)hipFreeAsync(ptr, st);
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
  int rc = ccs::noalloc_scratch_setup();
  if (rc) return rc;
  ccs_handle_s* h = new ccs_handle_s();
  h->stream = (hipStream_t)hip_stream;
  *handle = h;
  return CCS_OK;
}

extern "C" int ccs_destroy(ccs_handle_t h) {
  if (!h) return CCS_OK;
  for (ccs::Scratch* s : {&h->conv_w, &h->crf_a, &h->crf_msg, &h->crf_small, &h->crf_sums, &h->misc, &h->misc2, &h->crf_nan, &h->conv_y})
    s->release(h->stream);
  (void)hipStreamSynchronize(h->stream);
  delete h;
  return CCS_OK;
}

extern "C" int ccs_set_stream(ccs_handle_t h, void* hip_stream) {
  CCS_CHECK(h != nullptr, CCS_ERR_ARG, "ccs_set_stream: NULL handle");
  // Work queued on the old stream may still use the scratch, which from now on
  // is released and reused in the order of the new stream.
  if ((hipStream_t)hip_stream != h->stream) CCS_HIP(hipStreamSynchronize(h->stream));
  h->stream = (hipStream_t)hip_stream;
  return CCS_OK;
}

extern "C" int ccs_synchronize(ccs_handle_t h) {
  CCS_CHECK(h != nullptr, CCS_ERR_ARG, "ccs_synchronize: NULL handle");
  CCS_HIP(hipStreamSynchronize(h->stream));
  return CCS_OK;
}

extern "C" int ccs_reserve(ccs_handle_t h, int H, int W, int D, int C, int iters) {
  CCS_CHECK(h != nullptr, CCS_ERR_ARG, "ccs_reserve: NULL handle");
  CCS_CHECK(H >= 1 && W >= 1 && D >= 1 && C >= 1 && iters >= 0, CCS_ERR_ARG, "ccs_reserve: bad size");
  const size_t hw = (size_t)H * W, vol = hw * D;
  const size_t big = (size_t)(H > W ? H : W);
  auto mx = [](size_t a, size_t b) { return a > b ? a : b; };
  // Packed weights of the C -> C and 64 -> 64 layers, either precision.
  size_t packed = 0;
  for (int p : {CCS_FP32, CCS_BF16})
    for (int k : {C, 64}) packed = mx(packed, ccs_conv_packed_bytes(k, k, 3, 3, p));
  // misc: disparity list (correlation statics), image moments (up to 5
  // channels), labels + bound of the synchronising SlackProp statics.
  const size_t misc = mx(mx(sizeof(int) * (size_t)D, sizeof(double) * (2 * 5 * (size_t)ccs::kMomentBlocks + 10)),
                         sizeof(int) * hw + 2 * sizeof(double));
  // misc2: Gaussian temporary of the edge weights, float64 reduction partials.
  const size_t misc2 = mx(mx(sizeof(float) * hw, 2 * sizeof(double) * (size_t)ccs::kReduceBlocks),
                          sizeof(double) * (size_t)ccs::kXentBlocks);
  int rc;
  if (packed && (rc = h->conv_w.reserve(packed, h->stream))) return rc;
  if ((rc = h->misc.reserve(misc, h->stream))) return rc;
  if ((rc = h->misc2.reserve(misc2, h->stream))) return rc;
  if ((rc = h->crf_a.reserve(vol * sizeof(float), h->stream))) return rc;
  if ((rc = h->crf_msg.reserve(vol * sizeof(float), h->stream))) return rc;
  if ((rc = h->crf_small.reserve(sizeof(double) * 2 * big, h->stream))) return rc;
  if ((rc = h->crf_nan.reserve(sizeof(int) * (size_t)H, h->stream))) return rc;
  if ((rc = h->crf_sums.reserve(sizeof(double) * (4 * (size_t)iters + 2), h->stream))) return rc;
  return CCS_OK;
}
