// Handle, tensor and error plumbing of the C++ API (include/ccs_stereo.hpp).
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/ccs_stereo.hpp"

namespace ccs {

void check(int rc, const char* what) {
  if (rc != CCS_OK) throw Error(rc, std::string(what) + ": " + ccs_last_error());
}

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(CCS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t TensorDesc::elements() const {
  if (layout == Layout::NHWC) return (size_t)n * h * w * cstride();
  return (size_t)n * c * h * w;
}

size_t TensorDesc::bytes() const { return elements() * (precision == Precision::BF16 ? 2 : 4); }

bool TensorDesc::operator==(const TensorDesc& o) const {
  return n == o.n && c == o.c && h == o.h && w == o.w && layout == o.layout && cstride() == o.cstride() &&
         precision == o.precision;
}

Handle::Handle(void* hip_stream) : stream_(hip_stream) { check(ccs_create(&h_, hip_stream), "ccs_create"); }

Handle::~Handle() {
  if (h_) ccs_destroy(h_);
}

void Handle::setStream(void* hip_stream) {
  check(ccs_set_stream(h_, hip_stream), "ccs_set_stream");
  stream_ = hip_stream;
}

void Handle::synchronize() { check(ccs_synchronize(h_), "ccs_synchronize"); }

Tensor::Tensor(const TensorDesc& d) : desc_(d), owned_(true) {
  if (d.n < 1 || d.c < 1 || d.h < 1 || d.w < 1 || d.cstride() < d.c)
    throw Error(CCS_ERR_ARG, "Tensor: invalid shape");
  hip_check(hipMalloc(&ptr_, desc_.bytes()), "hipMalloc(Tensor)");
}

Tensor::Tensor(int n, int c, int h, int w, Layout layout) : Tensor(TensorDesc(n, c, h, w, layout)) {}

Tensor::Tensor(void* external, const TensorDesc& d) : desc_(d), ptr_(external), owned_(false) {
  if (!external) throw Error(CCS_ERR_ARG, "Tensor: external pointer is NULL");
}

Tensor::~Tensor() {
  if (owned_ && ptr_) (void)hipFree(ptr_);
}

void Tensor::copyFromHost(const void* host) {
  hip_check(hipMemcpy(ptr_, host, desc_.bytes(), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void Tensor::copyToHost(void* host) const {
  hip_check(hipMemcpy(host, ptr_, desc_.bytes(), hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

std::vector<float> Tensor::toNCHW(Handle& handle) const {
  const TensorDesc& d = desc_;
  std::vector<float> out((size_t)d.n * d.c * d.h * d.w);
  if (d.layout == Layout::NCHW && d.precision == Precision::FP32) {
    handle.synchronize();
    copyToHost(out.data());
    return out;
  }
  if (d.layout != Layout::NHWC) throw Error(CCS_ERR_UNSUPPORTED, "Tensor::toNCHW: bf16 planar tensors");
  float* tmp = nullptr;
  const size_t per = (size_t)d.c * d.h * d.w;
  hip_check(hipMalloc(&tmp, per * sizeof(float)), "hipMalloc(toNCHW)");
  const size_t esz = d.precision == Precision::BF16 ? 2 : 4;
  for (int i = 0; i < d.n; ++i) {
    const char* src = static_cast<const char*>(ptr_) + (size_t)i * d.h * d.w * d.cstride() * esz;
    int rc = ccs_nhwc_to_nchw(handle.get(), src, d.precision == Precision::BF16 ? CCS_BF16 : CCS_FP32, d.c,
                              d.cstride(), d.h, d.w, tmp);
    if (rc == CCS_OK) rc = ccs_synchronize(handle.get());
    if (rc != CCS_OK) {
      (void)hipFree(tmp);
      check(rc, "Tensor::toNCHW");
    }
    hipError_t e = hipMemcpy(out.data() + i * per, tmp, per * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      (void)hipFree(tmp);
      hip_check(e, "hipMemcpy D2H");
    }
  }
  (void)hipFree(tmp);
  return out;
}

size_t HostArray::size() const {
  size_t s = 1;
  for (int d : shape) s *= (size_t)d;
  return s;  // 1 for a 0-d array
}

}  // namespace ccs
