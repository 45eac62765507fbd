// Operator classes of the C++ API: the reference's operator layer
// (stereo/operators/*.cpp) over the HIP kernels of the C ABI.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ccs_stereo.hpp"

namespace ccs {
namespace {

int prec_code(Precision p) { return p == Precision::BF16 ? CCS_BF16 : CCS_FP32; }

Handle& default_handle() {
  thread_local std::unique_ptr<Handle> h;
  if (!h) h.reset(new Handle(nullptr));
  return *h;
}

[[noreturn]] void arg_fail(const std::string& msg) { throw Error(CCS_ERR_ARG, msg); }

int as_int_disp(float v, const char* what) {
  const float r = std::nearbyint(v);
  if (std::fabs(v - r) > 1e-6f)
    throw Error(CCS_ERR_UNSUPPORTED, std::string("StereoCorrelation: ") + what + " must be integer-valued");
  return (int)r;
}

}  // namespace

// ---------------------------------------------------------------- Operator

Operator::Operator(const TensorDesc& inTensorDesc, Type type) : m_inTensorDesc{inTensorDesc}, m_type(type) {}

Operator::Operator(int in, int ic, int ih, int iw, Type type)
    : m_inTensorDesc{TensorDesc(in, ic, ih, iw)}, m_type(type) {}

Operator::Operator(const std::vector<TensorDesc>& inTensorDesc, Type type)
    : m_inTensorDesc(inTensorDesc), m_type(type) {
  if (m_inTensorDesc.empty()) arg_fail("Operator: no input descriptors");
}

Operator::~Operator() = default;

Tensor* Operator::forward(Tensor* d_input, Handle& handle) {
  std::vector<Tensor*> in{d_input};
  return forward(in, handle);
}

void Operator::setParams(const std::string& npzPath, const std::string& array) {
  setParams(npz::load_array(npzPath, array));
}

void Operator::checkParams(const HostArray& h_params, const TensorDesc& d) {
  const size_t want = (size_t)d.n * d.c * d.h * d.w;
  if (h_params.shape.size() != 1 && h_params.shape.size() != 4)
    arg_fail("setParams: only 1-D and 4-D parameter arrays are accepted");
  if (h_params.size() != want || h_params.data.size() != want)
    arg_fail("setParams: parameter has " + std::to_string(h_params.size()) + " elements, operator needs " +
             std::to_string(want));
  if (h_params.shape.size() == 4 &&
      (h_params.shape[0] != d.n || h_params.shape[1] != d.c || h_params.shape[2] != d.h || h_params.shape[3] != d.w))
    arg_fail("setParams: 4-D parameter shape does not match the operator");
}

void Operator::setParams(const HostArray& h_params) {
  Tensor* p = getParams();
  if (!p) arg_fail("setParams: operator has no parameters");
  checkParams(h_params, p->desc());
  p->copyFromHost(h_params.data.data());
  paramsChanged();
}

// ------------------------------------------------------------- Convolution

Convolution::Convolution(Mode convMode, const TensorDesc& inTensorDesc, int fk, int fc, int fh, int fw,
                         Tensor* d_out, int pad_x, int pad_y)
    : Operator(inTensorDesc, Operator::CONVOLUTION), m_convMode(convMode), m_fk(fk), m_fc(fc), m_fh(fh), m_fw(fw) {
  initialize(d_out, pad_x, pad_y);
}

Convolution::Convolution(Mode convMode, int in, int ic, int ih, int iw, int fk, int fc, int fh, int fw,
                         Tensor* d_out, int pad_x, int pad_y)
    : Operator(in, ic, ih, iw, Operator::CONVOLUTION), m_convMode(convMode), m_fk(fk), m_fc(fc), m_fh(fh), m_fw(fw) {
  initialize(d_out, pad_x, pad_y);
}

void Convolution::initialize(Tensor* d_out, int pad_x, int pad_y) {
  const TensorDesc& in = m_inTensorDesc[0];
  if (pad_x != 0 || pad_y != 0)
    throw Error(CCS_ERR_UNSUPPORTED, "Convolution: only pad 0 is implemented (the networks use valid convolutions)");
  if (in.n != 1) throw Error(CCS_ERR_UNSUPPORTED, "Convolution: batch size 1 only");
  if (m_fc != in.c) arg_fail("Convolution: filter channels do not match the input");
  if (m_fh > in.h || m_fw > in.w || m_fk < 1) arg_fail("Convolution: filter larger than the input");
  m_filter = new Tensor(TensorDesc(m_fk, m_fc, m_fh, m_fw));
  allocateOutputMemory(d_out);
}

Convolution::~Convolution() {
  deleteOutputMemory();
  delete m_filter;
  delete m_packed;
}

TensorDesc Convolution::outTensorDesc() {
  const TensorDesc& in = m_inTensorDesc[0];
  return TensorDesc(1, m_fk, in.h - m_fh + 1, in.w - m_fw + 1, Layout::NHWC, ccs_feature_stride(m_fk), m_precision);
}

void Convolution::allocateOutputMemory(Tensor* d_out) {
  if (d_out) {
    if (!(d_out->desc() == outTensorDesc())) arg_fail("Convolution: external output has the wrong shape/layout");
    m_output = d_out;
    m_extOutPtr = true;
  } else if (!m_output) {
    m_output = new Tensor(outTensorDesc());
    m_extOutPtr = false;
  }
}

void Convolution::deleteOutputMemory() {
  if (m_output && !m_extOutPtr) delete m_output;
  if (!m_extOutPtr) m_output = nullptr;
}

void Convolution::setPrecision(Precision p) {
  if (p == m_precision) return;
  if (m_extOutPtr) arg_fail("Convolution::setPrecision: the output buffer is external");
  m_precision = p;
  deleteOutputMemory();
  allocateOutputMemory(nullptr);
}

void Convolution::setFusedEpilogue(const Tensor* bias, bool tanh) {
  m_fusedBias = bias;
  m_fusedTanh = tanh;
}

Tensor* Convolution::forward(std::vector<Tensor*>& d_inputs, Handle& handle) {
  if (d_inputs.empty() || !d_inputs[0]) arg_fail("Convolution::forward: no input");
  const Tensor& x = *d_inputs[0];
  const TensorDesc& xd = x.desc();
  const TensorDesc& in = m_inTensorDesc[0];
  if (xd.n != 1 || xd.c != in.c || xd.h != in.h || xd.w != in.w)
    arg_fail("Convolution::forward: input shape differs from the constructor's");
  int layout;
  if (xd.layout == Layout::NCHW) {
    if (xd.precision != Precision::FP32 || xd.c > 8)
      throw Error(CCS_ERR_UNSUPPORTED, "Convolution::forward: planar input must be a float32 image (C <= 8)");
    layout = CCS_LAYOUT_NCHW;
  } else {
    if (xd.cstride() != ccs_feature_stride(xd.c) || xd.precision != m_precision)
      arg_fail("Convolution::forward: channel-last input needs stride round_up(C, 8) in the operator's precision");
    layout = CCS_LAYOUT_NHWC;
  }
  allocateOutputMemory(nullptr);
  const int prec = prec_code(m_precision);
  const float* b = m_fusedBias ? m_fusedBias->data() : nullptr;
  const int act = m_fusedTanh ? 1 : 0;
  // Weights in the orientation of a true convolution (the kernels flip them).
  const float* w = m_filter->data();
  std::unique_ptr<Tensor> flipped;
  if (m_convMode == CROSS_CORRELATION) {
    std::vector<float> hw((size_t)m_fk * m_fc * m_fh * m_fw), hf(hw.size());
    m_filter->copyToHost(hw.data());
    for (int k = 0; k < m_fk * m_fc; ++k)
      for (int i = 0; i < m_fh; ++i)
        for (int j = 0; j < m_fw; ++j)
          hf[((size_t)k * m_fh + i) * m_fw + j] = hw[((size_t)k * m_fh + (m_fh - 1 - i)) * m_fw + (m_fw - 1 - j)];
    flipped.reset(new Tensor(m_filter->desc()));
    flipped->copyFromHost(hf.data());
    w = flipped->data();
  }
  if (layout == CCS_LAYOUT_NHWC && !flipped) {
    // Packed MFMA path: pack once per parameter upload / precision.
    if (!m_packedValid || m_packedPrecision != m_precision) {
      delete m_packed;
      m_packed = nullptr;
      const size_t bytes = ccs_conv_packed_bytes(m_fk, m_fc, m_fh, m_fw, prec);
      m_packed = new Tensor(TensorDesc(1, 1, 1, (int)((bytes + 3) / 4)));
      check(ccs_conv_pack_weights(handle.get(), w, m_fk, m_fc, m_fh, m_fw, prec, m_packed->raw()),
            "Convolution: pack weights");
      m_packedValid = true;
      m_packedPrecision = m_precision;
    }
    const int rc = ccs_conv_forward_packed(handle.get(), x.raw(), layout, xd.h, xd.w, xd.c, m_packed->raw(), b,
                                           m_fk, m_fh, m_fw, act, prec, m_output->raw());
    if (rc == CCS_OK) return m_output;
    if (rc != CCS_ERR_UNSUPPORTED) check(rc, "Convolution::forward");
  }
  check(ccs_conv2d_forward(handle.get(), x.raw(), layout, xd.h, xd.w, xd.c, w, b, m_fk, m_fh, m_fw, act, prec,
                           m_output->raw()),
        "Convolution::forward");
  if (flipped) handle.synchronize();  // the temporary filter must outlive the launch
  return m_output;
}

// -------------------------------------------------------------------- Bias

Bias::Bias(const TensorDesc& inTensorDesc, const TensorDesc& biasTensorDesc) : Operator(inTensorDesc, BIAS) {
  if ((size_t)biasTensorDesc.n * biasTensorDesc.c * biasTensorDesc.h * biasTensorDesc.w != (size_t)inTensorDesc.c)
    arg_fail("Bias: one value per channel");
  m_bias = new Tensor(TensorDesc(biasTensorDesc.n, biasTensorDesc.c, biasTensorDesc.h, biasTensorDesc.w));
}

Bias::Bias(const TensorDesc& inTensorDesc, int fk, int fc, int fh, int fw)
    : Bias(inTensorDesc, TensorDesc(fk, fc, fh, fw)) {}

Bias::~Bias() { delete m_bias; }

Tensor* Bias::forward(std::vector<Tensor*>& d_inputs, Handle& handle) {
  if (d_inputs.empty() || !d_inputs[0]) arg_fail("Bias::forward: no input");
  Tensor& y = *d_inputs[0];
  const TensorDesc& d = y.desc();
  if (d.precision != Precision::FP32 || d.n != 1) throw Error(CCS_ERR_UNSUPPORTED, "Bias::forward: float32, N=1");
  check(ccs_bias_forward(handle.get(), y.data(), d.layout == Layout::NHWC ? CCS_LAYOUT_NHWC : CCS_LAYOUT_NCHW, d.c,
                         d.h, d.w, d.cstride(), m_bias->data()),
        "Bias::forward");
  return &y;
}

// -------------------------------------------------------------- Activation

Activation::Activation(const TensorDesc& inOutTensorDesc, Mode, double) : Operator(inOutTensorDesc, ACTIVATION) {}

Tensor* Activation::forward(std::vector<Tensor*>& d_inputs, Handle& handle) {
  if (d_inputs.empty() || !d_inputs[0]) arg_fail("Activation::forward: no input");
  Tensor& y = *d_inputs[0];
  if (y.desc().precision != Precision::FP32) throw Error(CCS_ERR_UNSUPPORTED, "Activation::forward: float32 only");
  check(ccs_tanh_forward(handle.get(), y.data(), (long)y.desc().elements()), "Activation::forward");
  return &y;
}

// ----------------------------------------------------------------- Softmax

Softmax::Softmax(const TensorDesc& inOutTensorDesc, Mode) : Operator(inOutTensorDesc, SOFTMAX) {}

Tensor* Softmax::forward(std::vector<Tensor*>& d_inputs, Handle& handle) {
  if (d_inputs.empty() || !d_inputs[0]) arg_fail("Softmax::forward: no input");
  Tensor& x = *d_inputs[0];
  const TensorDesc& d = x.desc();
  if (d.layout != Layout::NHWC || d.cstride() != d.c || d.precision != Precision::FP32)
    throw Error(CCS_ERR_UNSUPPORTED, "Softmax::forward: dense NHWC float32 volumes only");
  check(ccs_softmax_forward(handle.get(), x.data(), (long)d.n * d.h * d.w, d.c, m_negate ? 1 : 0),
        "Softmax::forward");
  return &x;
}

// ------------------------------------------------------- StereoCorrelation

StereoCorrelation::StereoCorrelation(const std::vector<TensorDesc>& inTensorDesc, Tensor* d_out, float min_disp,
                                     float max_disp, float step, int rect_corr)
    : Operator(inTensorDesc, STEREO_CORRELATION), m_rectCorr(rect_corr) {
  if (m_inTensorDesc.size() != 2) arg_fail("StereoCorrelation: two inputs");
  const TensorDesc &a = m_inTensorDesc[0], &b = m_inTensorDesc[1];
  if (a.c != b.c || a.h != b.h || a.w != b.w || a.n != 1 || b.n != 1)
    arg_fail("StereoCorrelation: inputs must have equal shapes, N=1");
  if (step == 0) {  // reference default: offsets 0..127 (stereo/operators/stereocorrelation.cpp:36)
    min_disp = 0;
    max_disp = 127;
    step = 1;
  }
  setDisparities(min_disp, max_disp, step, d_out);
}

StereoCorrelation::~StereoCorrelation() {
  if (!m_extOutPtr) delete m_output;
}

TensorDesc StereoCorrelation::outTensorDesc() {
  const TensorDesc& a = m_inTensorDesc[0];
  return TensorDesc(1, m_numDisps, a.h, a.w, Layout::NHWC);
}

void StereoCorrelation::setDisparities(float min_disp, float max_disp, float step, Tensor* d_out) {
  if (step <= 0 || max_disp < min_disp) arg_fail("StereoCorrelation::setDisparities: need step > 0, max >= min");
  m_dispMin = as_int_disp(min_disp, "min_disp");
  m_dispStep = as_int_disp(step, "step");
  int n = 0;
  for (float d = min_disp; d <= max_disp; d += step) ++n;  // the reference's float loop (:83-85)
  if (n > 256) throw Error(CCS_ERR_UNSUPPORTED, "StereoCorrelation: at most 256 disparities");
  m_numDisps = n;
  allocateOutput(d_out);
}

void StereoCorrelation::allocateOutput(Tensor* d_out) {
  const TensorDesc want = outTensorDesc();
  if (d_out) {
    const TensorDesc& g = d_out->desc();
    // The reference checks NHWC and matching dimensions (stereocorrelation.cpp:101-133).
    if (g.layout != Layout::NHWC || g.n != 1 || g.c != want.c || g.h != want.h || g.w != want.w ||
        g.precision != Precision::FP32 || g.cstride() != g.c)
      arg_fail("StereoCorrelation: external output must be NHWC 1 x D x H x W float32");
    if (m_output && !m_extOutPtr) delete m_output;
    m_output = d_out;
    m_extOutPtr = true;
  } else if (!m_output || m_extOutPtr || !(m_output->desc() == want)) {
    if (m_output && !m_extOutPtr) delete m_output;
    m_output = new Tensor(want);
    m_extOutPtr = false;
  }
}

Tensor* StereoCorrelation::forward(std::vector<Tensor*>& d_inputs, Handle& handle) {
  if (d_inputs.size() < 2 || !d_inputs[0] || !d_inputs[1]) arg_fail("StereoCorrelation::forward: two inputs");
  const Tensor &f0 = *d_inputs[0], &f1 = *d_inputs[1];
  const TensorDesc& d = f0.desc();
  if (!(d == f1.desc())) arg_fail("StereoCorrelation::forward: inputs differ");
  if (d.c != m_inTensorDesc[0].c || d.h != m_inTensorDesc[0].h || d.w != m_inTensorDesc[0].w)
    arg_fail("StereoCorrelation::forward: input shape differs from the constructor's");
  if (d.layout == Layout::NHWC && m_rectCorr == 0) {
    if (d.cstride() != ccs_feature_stride(d.c))
      arg_fail("StereoCorrelation::forward: channel-last inputs need stride round_up(C, 8)");
    check(ccs_correlation_volume(handle.get(), f0.raw(), f1.raw(), prec_code(d.precision), d.h, d.w, d.c, m_dispMin,
                                 m_dispStep, m_numDisps, m_negSoftmax ? CCS_EPILOGUE_NEG_SOFTMAX : CCS_EPILOGUE_RAW,
                                 m_output->data()),
          "StereoCorrelation::forward");
    return m_output;
  }
  // Rectification correction or planar inputs: the reference's kernel form.
  std::vector<float> disps(m_numDisps);
  for (int k = 0; k < m_numDisps; ++k) disps[k] = -(float)(m_dispMin + k * m_dispStep);
  Tensor dd(TensorDesc(1, 1, 1, m_numDisps));
  dd.copyFromHost(disps.data());
  std::unique_ptr<Tensor> p0, p1;
  const float *a = f0.data(), *b = f1.data();
  if (d.layout == Layout::NHWC) {
    p0.reset(new Tensor(TensorDesc(1, d.c, d.h, d.w)));
    p1.reset(new Tensor(TensorDesc(1, d.c, d.h, d.w)));
    check(ccs_nhwc_to_nchw(handle.get(), f0.raw(), prec_code(d.precision), d.c, d.cstride(), d.h, d.w, p0->data()),
          "StereoCorrelation: layout");
    check(ccs_nhwc_to_nchw(handle.get(), f1.raw(), prec_code(d.precision), d.c, d.cstride(), d.h, d.w, p1->data()),
          "StereoCorrelation: layout");
    a = p0->data();
    b = p1->data();
  } else if (d.precision != Precision::FP32) {
    throw Error(CCS_ERR_UNSUPPORTED, "StereoCorrelation: planar inputs must be float32");
  }
  check(ccs_correlation_forward(handle.get(), a, b, dd.data(), m_output->data(), 1, d.c, d.h, d.w, m_numDisps,
                                CCS_LAYOUT_NHWC, m_rectCorr),
        "StereoCorrelation::forward");
  if (m_negSoftmax)
    check(ccs_softmax_forward(handle.get(), m_output->data(), (long)d.h * d.w, m_numDisps, 1),
          "StereoCorrelation: softmax");
  handle.synchronize();  // temporaries above are freed on return
  return m_output;
}

void StereoCorrelation::forward(float* input0, float* input1, float* disparities, float* output, int in, int ic,
                                int ih, int iw, int lenDisps, Layout memoryLayout, int rectCorr, Handle* handle) {
  Handle& h = handle ? *handle : default_handle();
  check(ccs_correlation_forward(h.get(), input0, input1, disparities, output, in, ic, ih, iw, lenDisps,
                                memoryLayout == Layout::NHWC ? CCS_LAYOUT_NHWC : CCS_LAYOUT_NCHW, rectCorr),
        "StereoCorrelation::forward");
}

void StereoCorrelation::backward(float* d_outGrad0, float* d_outGrad1, float* d_inGrad, float* d_im0, float* d_im1,
                                 float* d_disparities, int n, int c, int h, int w, int numDisps, Layout memoryLayout,
                                 Handle* handle) {
  Handle& hd = handle ? *handle : default_handle();
  check(ccs_correlation_backward(hd.get(), d_outGrad0, d_outGrad1, d_inGrad, d_im0, d_im1, d_disparities, n, c, h, w,
                                 numDisps, memoryLayout == Layout::NHWC ? CCS_LAYOUT_NHWC : CCS_LAYOUT_NCHW),
        "StereoCorrelation::backward");
}

// --------------------------------------------------------------- SlackProp

SlackProp::SlackProp(const TensorDesc& inTensorDesc) : Operator(inTensorDesc, SLACKPROP) {}

SlackProp::~SlackProp() { delete m_output; }

TensorDesc SlackProp::outTensorDesc() {
  const TensorDesc& in = m_inTensorDesc[0];
  return TensorDesc(1, 1, in.h, in.w);
}

Tensor* SlackProp::forward(std::vector<Tensor*>& d_inputs, Handle& handle) {
  if (d_inputs.size() < 3 || !d_inputs[0] || !d_inputs[1] || !d_inputs[2])
    arg_fail("SlackProp::forward: inputs {cost volume, wx, wy}");
  const Tensor& C = *d_inputs[0];
  const TensorDesc& d = C.desc();
  if (d.layout != Layout::NHWC || d.n != 1 || d.cstride() != d.c || d.precision != Precision::FP32)
    arg_fail("SlackProp::forward: the cost volume must be NHWC 1 x D x H x W float32");
  if (!m_output) m_output = new Tensor(outTensorDesc());
  if (m_output->height() != d.h || m_output->width() != d.w) arg_fail("SlackProp::forward: shape changed");
  Tensor labels(TensorDesc(1, 1, d.h, d.w));  // int32 storage
  Tensor lb(TensorDesc(1, 1, 1, 2));          // one double
  check(ccs_slackprop(handle.get(), C.data(), d_inputs[1]->data(), d_inputs[2]->data(), d.h, d.w, d.c, m_P1, m_P2,
                      m_delta, m_iters, (int*)labels.raw(), nullptr, nullptr, (double*)lb.raw()),
        "SlackProp::forward");
  check(ccs_remap(handle.get(), (const int*)labels.raw(), d.h, d.w, 0.f, 1.f, m_output->data()),
        "SlackProp::forward: labels");
  check(ccs_slackprop_status(handle.get()), "SlackProp::forward");  // synchronises; NaN costs throw
  double v = 0;
  if (hipMemcpy(&v, lb.raw(), sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    throw Error(CCS_ERR_HIP, "SlackProp::forward: reading the bound");
  m_lowerBound = (float)v;
  return m_output;
}

float SlackProp::forward(Tensor* costVolume, Tensor* pwVolume, Tensor* output, int max_iter, Handle* handle) {
  if (!costVolume || !pwVolume || !output) arg_fail("SlackProp::forward: NULL tensor");
  const TensorDesc& c = costVolume->desc();
  const TensorDesc& p = pwVolume->desc();
  const TensorDesc& o = output->desc();
  // attention! NHWC cost volume only, as in the reference (slackprop.cpp:70)
  if (c.layout != Layout::NHWC || c.n != 1 || c.cstride() != c.c || c.precision != Precision::FP32)
    arg_fail("SlackProp::forward: the cost volume must be NHWC 1 x D x H x W float32");
  if (p.layout != Layout::NCHW || p.n != 1 || p.h != c.h || p.w != c.w || p.precision != Precision::FP32)
    arg_fail("SlackProp::forward: the pairwise volume must be NCHW 1 x 2 levels x H x W float32");
  if (o.n != 1 || o.c != 1 || o.h != c.h || o.w != c.w || o.precision != Precision::FP32)
    arg_fail("SlackProp::forward: the output must be 1 x 1 x H x W float32");
  Handle& h = handle ? *handle : default_handle();
  float lb = 0.f;
  check(ccs_slackprop_forward_pw(h.get(), costVolume->data(), pwVolume->data(), output->data(), 1, c.c, c.h, c.w, p.c,
                                 max_iter, &lb),
        "SlackProp::forward");
  return lb;
}

float SlackProp::forward(float* costVolume, float* wx, float* wy, float* output, int in, int ic, int ih, int iw,
                         float P1, float P2, int max_iter, Handle* handle) {
  Handle& h = handle ? *handle : default_handle();
  float lb = 0.f;
  check(ccs_slackprop_forward(h.get(), costVolume, wx, wy, output, in, ic, ih, iw, P1, P2, max_iter, &lb),
        "SlackProp::forward");
  return lb;
}

float SlackProp::forwardVolumeOut(float* costVolume, float* wx, float* wy, float* output, int in, int ic, int ih,
                                  int iw, float P1, float P2, int max_iter, int delta, Handle* handle) {
  Handle& h = handle ? *handle : default_handle();
  float lb = 0.f;
  check(ccs_slackprop_forward_volume_out(h.get(), costVolume, wx, wy, output, in, ic, ih, iw, P1, P2, max_iter, delta,
                                         &lb),
        "SlackProp::forwardVolumeOut");
  return lb;
}

}  // namespace ccs
