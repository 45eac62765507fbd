// StereoNet / ColorStereoNet (stereo/stereonet.cpp, stereo/colorstereonet.cpp)
// and the Stereo facade (stereo/stereo.cpp) over the operator classes.
//
// The network keeps the reference's operator lists (Convolution, Bias,
// Activation per layer; StereoCorrelation, Softmax after the towers) so that
// setAllParams walks the same arr_<i> keys, but it runs them fused: each
// Convolution applies its layer's bias and tanh in the MFMA epilogue, and the
// correlation applies the -softmax over D in its epilogue.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ccs_stereo.hpp"

namespace ccs {
namespace {

[[noreturn]] void arg_fail(const std::string& msg) { throw Error(CCS_ERR_ARG, msg); }

int prec_code(Precision p) { return p == Precision::BF16 ? CCS_BF16 : CCS_FP32; }

// One convolution of the networks: K filters over C channels, KH x KW, with or
// without a bias (and tanh) after it.
struct LayerSpec {
  int K, C, KH, KW;
  bool bias;
};

// Tower: layer 1 3x3, ic -> 100; layers 2..numLayers 2x2, 100 -> 100
// (stereo/stereonet.cpp:65-132).
std::vector<LayerSpec> tower_spec(int numLayers, int ic) {
  std::vector<LayerSpec> v;
  for (int l = 0; l < numLayers; ++l) v.push_back(l == 0 ? LayerSpec{100, ic, 3, 3, true} : LayerSpec{100, 100, 2, 2, true});
  return v;
}

// Pairwise network on the RGB input (stereo/stereonet.cpp:143-168): 3x3 3 -> 64,
// 3x3 64 -> 64, each with bias + tanh, then the 1x1 64 -> 2 head without bias.
std::vector<LayerSpec> pairwise_spec() { return {{64, 3, 3, 3, true}, {64, 64, 3, 3, true}, {2, 64, 1, 1, false}}; }

int log2i(int num) {  // floor(log2(num)) for num >= 1, 0 for num <= 1 (stereo/stereo.cpp:56-62)
  int r = 0;
  while (num >>= 1) ++r;
  return r;
}

}  // namespace

// ----------------------------------------------------------------- StereoNet

StereoNet::StereoNet(int numLayers, int in, int ic, int ih, int iw)
    : m_handle(new Handle(nullptr)), m_numLayers(numLayers), m_in(in), m_ic(ic), m_ih(ih), m_iw(iw) {
  if (numLayers != 3 && numLayers != 5 && numLayers != 7)
    arg_fail("StereoNet: numLayers must be 3, 5 or 7 (stereo/stereo.cpp:198-202)");
  if (in != 1) throw Error(CCS_ERR_UNSUPPORTED, "StereoNet: batch size 1 only");
  if (ic < 1 || ic > 5) throw Error(CCS_ERR_UNSUPPORTED, "StereoNet: 1 to 5 input channels");
}

StereoNet::~StereoNet() {
  if (m_stream2) (void)hipStreamSynchronize((hipStream_t)m_stream2);
  for (auto* v : {&m_leftOps, &m_rightOps, &m_lrOps, &m_pairwiseOps})
    for (Operator* op : *v) delete op;
  m_handle2.reset();
  if (m_evIn) (void)hipEventDestroy((hipEvent_t)m_evIn);
  if (m_evOut) (void)hipEventDestroy((hipEvent_t)m_evOut);
  if (m_stream2) (void)hipStreamDestroy((hipStream_t)m_stream2);
}

void StereoNet::initNet(float min_disp, float max_disp, float step, int rect_corr, Tensor* d_unaryOut,
                        Tensor* d_pairwiseOut) {
  if (m_initialized) arg_fail("StereoNet::initNet: already initialised");
  const std::vector<LayerSpec> spec = tower_spec(m_numLayers, m_ic);
  for (std::vector<Operator*>* ops : {&m_leftOps, &m_rightOps}) {
    for (size_t l = 0; l < spec.size(); ++l) {
      const LayerSpec& L = spec[l];
      Convolution* conv = l == 0 ? new Convolution(Convolution::CONVOLUTION, m_in, m_ic, m_ih, m_iw, L.K, L.C, L.KH, L.KW)
                                 : new Convolution(Convolution::CONVOLUTION, ops->back()->outTensorDesc(), L.K, L.C, L.KH, L.KW);
      conv->setPrecision(m_precision);
      ops->push_back(conv);
      Bias* bias = new Bias(conv->outTensorDesc(), 1, L.K, 1, 1);
      ops->push_back(bias);
      ops->push_back(new Activation(conv->outTensorDesc(), Activation::TANH));
      conv->setFusedEpilogue(bias->getParams(), true);
    }
  }
  std::vector<TensorDesc> corrInputs = {m_leftOps.back()->outTensorDesc(), m_rightOps.back()->outTensorDesc()};
  auto* corr = new StereoCorrelation(corrInputs, d_unaryOut, min_disp, max_disp, step, rect_corr);
  corr->setNegSoftmaxEpilogue(true);
  m_lrOps.push_back(corr);
  auto* sm = new Softmax(corr->outTensorDesc(), Softmax::CHANNEL);
  sm->setNegate(true);
  m_lrOps.push_back(sm);
  if (d_pairwiseOut) initPairwise(d_pairwiseOut);
  m_initialized = true;
}

// Pairwise network (stereo/stereonet.cpp:143-168): on the colour input, conv
// 3x3 3->64 + bias + tanh, conv 3x3 64->64 + bias + tanh, conv 1x1 64->2
// without bias, then |.| (:402) into d_pairwiseOut = (wx, wy).  The towers
// shrink the padded input by numLayers + 1 pixels and this net by 4, so it
// reads the input cropped by (numLayers - 3) / 2 pixels per side: 2 for seven
// layers as in stereo/colorstereonet.cpp:69-73, none for three.
void StereoNet::initPairwise(Tensor* d_pairwiseOut) {
  if (m_ic != 3 && m_ic != 5)
    throw Error(CCS_ERR_UNSUPPORTED, "StereoNet::initNet: the pairwise network takes the colour input");
  m_pwCrop = (m_numLayers - 3) / 2;
  const int ih = m_ih - 2 * m_pwCrop, iw = m_iw - 2 * m_pwCrop;
  const TensorDesc& od = d_pairwiseOut->desc();
  if (od.n != 1 || od.c != 2 || od.h != ih - 4 || od.w != iw - 4 || od.layout != Layout::NCHW ||
      od.precision != Precision::FP32)
    arg_fail("StereoNet::initNet: d_pairwiseOut must be a 1 x 2 x " + std::to_string(ih - 4) + " x " +
             std::to_string(iw - 4) + " float32 planar tensor");
  const std::vector<LayerSpec> spec = pairwise_spec();
  for (size_t l = 0; l < spec.size(); ++l) {
    const LayerSpec& L = spec[l];
    auto* c = l == 0 ? new Convolution(Convolution::CONVOLUTION, m_in, L.C, ih, iw, L.K, L.C, L.KH, L.KW)
                     : new Convolution(Convolution::CONVOLUTION, m_pairwiseOps.back()->outTensorDesc(), L.K, L.C,
                                       L.KH, L.KW);
    m_pairwiseOps.push_back(c);
    if (!L.bias) {
      c->deleteOutputMemory();  // the head kernel writes d_pairwiseOut directly
      continue;
    }
    c->setPrecision(m_precision);
    auto* b = new Bias(c->outTensorDesc(), 1, L.K, 1, 1);
    m_pairwiseOps.push_back(b);
    m_pairwiseOps.push_back(new Activation(c->outTensorDesc(), Activation::TANH));
    c->setFusedEpilogue(b->getParams(), true);
  }
  m_pairwiseOut = d_pairwiseOut;
}

void StereoNet::setPrecision(Precision p) {
  m_precision = p;
  for (auto* v : {&m_leftOps, &m_rightOps})
    for (Operator* op : *v)
      if (op->type() == Operator::CONVOLUTION) static_cast<Convolution*>(op)->setPrecision(p);
}

void StereoNet::setAllParams(const std::string& npzPath) {
  if (!m_initialized) arg_fail("StereoNet::setAllParams: call initNet first");
  const std::map<std::string, HostArray> all = npz::load(npzPath);
  int param_idx = 0;
  for (size_t i = 0; i < m_leftOps.size(); ++i) {
    if (m_leftOps[i]->type() == Operator::ACTIVATION) continue;
    const std::string key = "arr_" + std::to_string(param_idx);
    auto it = all.find(key);
    if (it == all.end()) throw Error(CCS_ERR_IO, npzPath + ": no array '" + key + "'");
    if (m_verbose) std::printf("Unary: params_idx %d\n", param_idx);
    // The same array goes into both towers (shared weights, stereo/stereonet.cpp:196-206).
    m_leftOps[i]->setParams(it->second);
    m_rightOps[i]->setParams(it->second);
    ++param_idx;
  }
  // Then the pairwise network's parameters, in the same walk (stereonet.cpp:209-224).
  for (Operator* op : m_pairwiseOps) {
    if (op->type() == Operator::ACTIVATION) continue;
    const std::string key = "arr_" + std::to_string(param_idx);
    auto it = all.find(key);
    if (it == all.end()) throw Error(CCS_ERR_IO, npzPath + ": no array '" + key + "' (pairwise network)");
    if (m_verbose) std::printf("PW: params_idx %d\n", param_idx);
    op->setParams(it->second);
    ++param_idx;
  }
  // Trailing arrays (the learned L1, L2 of CRF models) are not network parameters.
}

std::vector<TensorDesc> StereoNet::parameterShapes(int numLayers, int ic, bool pairwise) {
  if (numLayers != 3 && numLayers != 5 && numLayers != 7)
    arg_fail("StereoNet: numLayers must be 3, 5 or 7 (stereo/stereo.cpp:198-202)");
  std::vector<TensorDesc> v;
  auto add = [&](const std::vector<LayerSpec>& spec) {
    for (const LayerSpec& L : spec) {
      v.push_back(TensorDesc(L.K, L.C, L.KH, L.KW));
      if (L.bias) v.push_back(TensorDesc(1, L.K, 1, 1));
    }
  };
  add(tower_spec(numLayers, ic));
  if (pairwise) add(pairwise_spec());
  return v;
}

std::vector<std::string> StereoNet::checkParameterFile(const std::string& npzPath, int numLayers, int ic,
                                                       bool pairwise) {
  const std::map<std::string, HostArray> all = npz::load(npzPath);
  const std::vector<TensorDesc> shapes = parameterShapes(numLayers, ic, pairwise);
  for (size_t i = 0; i < shapes.size(); ++i) {
    const std::string key = "arr_" + std::to_string(i);
    auto it = all.find(key);
    if (it == all.end()) throw Error(CCS_ERR_IO, npzPath + ": no array '" + key + "'");
    try {
      Operator::checkParams(it->second, shapes[i]);
    } catch (const Error& e) {
      throw Error(e.status(), npzPath + ": " + key + ": " + e.what());
    }
  }
  std::vector<std::string> unread;
  for (size_t i = shapes.size();; ++i) {
    const std::string key = "arr_" + std::to_string(i);
    if (!all.count(key)) break;
    unread.push_back(key);
  }
  return unread;
}

void StereoNet::setDisparities(float min_disp, float max_disp, float step, Tensor* d_out) {
  if (m_lrOps.empty() || m_lrOps[0]->type() != Operator::STEREO_CORRELATION)
    arg_fail("StereoNet::setDisparities: no StereoCorrelation operation (call initNet first)");
  static_cast<StereoCorrelation*>(m_lrOps[0])->setDisparities(min_disp, max_disp, step, d_out);
}

Tensor* StereoNet::performPrediction(Tensor* d_inputLeft, Tensor* d_inputRight) {
  if (!m_initialized) arg_fail("StereoNet: call initNet first");
  Handle& h = *m_handle;
  // Without garbage collection the two towers run concurrently: the right one
  // on a second, non-blocking stream of default priority (each layer's tail
  // then overlaps the other tower's layer).  With it, every free synchronises
  // the device anyway, so one stream is used.
  const bool two = !m_allowGc;
  if (two && !m_stream2) {
    hipStream_t s2;
    hipEvent_t e0, e1;
    if (hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e0, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e1, hipEventDisableTiming) != hipSuccess)
      throw Error(CCS_ERR_HIP, "StereoNet: creating the second stream");
    m_stream2 = s2;
    m_evIn = e0;
    m_evOut = e1;
    m_handle2.reset(new Handle(s2));
  }
  auto hip_ok = [](hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(CCS_ERR_HIP, std::string("StereoNet: ") + what + ": " + hipGetErrorString(e));
  };
  if (two) {  // the right tower starts after everything queued so far on the main stream
    hip_ok(hipEventRecord((hipEvent_t)m_evIn, (hipStream_t)h.stream()), "event record");
    hip_ok(hipStreamWaitEvent((hipStream_t)m_stream2, (hipEvent_t)m_evIn, 0), "stream wait");
  }
  Tensor* feats[2] = {d_inputLeft, d_inputRight};
  for (int t = 1; t >= 0; --t) {  // right tower first, so it is queued on its stream before the left one
    std::vector<Operator*>& ops = t == 0 ? m_leftOps : m_rightOps;
    Handle& ht = (two && t == 1) ? *m_handle2 : h;
    Tensor* x = feats[t];
    Convolution* prev = nullptr;
    for (Operator* op : ops) {
      if (op->type() != Operator::CONVOLUTION) continue;  // bias + tanh run in the convolution epilogue
      auto* conv = static_cast<Convolution*>(op);
      x = conv->forward(x, ht);
      if (prev && m_allowGc) prev->deleteOutputMemory();  // freed once its consumer is queued
      prev = conv;
    }
    feats[t] = x;
  }
  if (two) {  // the correlation waits for the right tower
    hip_ok(hipEventRecord((hipEvent_t)m_evOut, (hipStream_t)m_stream2), "event record");
    hip_ok(hipStreamWaitEvent((hipStream_t)h.stream(), (hipEvent_t)m_evOut, 0), "stream wait");
  }
  std::vector<Tensor*> in{feats[0], feats[1]};
  Tensor* vol = m_lrOps[0]->forward(in, h);  // correlation + fused -softmax
  if (!m_pairwiseOps.empty()) runPairwise(d_inputLeft);
  if (m_allowGc) {
    h.synchronize();
    for (auto* v : {&m_leftOps, &m_rightOps})
      if (!v->empty())
        for (Operator* op : *v)
          if (op->type() == Operator::CONVOLUTION) static_cast<Convolution*>(op)->deleteOutputMemory();
  }
  return vol;
}

void StereoNet::runPairwise(Tensor* d_inputLeft) {
  Handle& h = *m_handle;
  // The pairwise net reads the RGB planes only (with the x/y model the input
  // has two more planes after them, stereo/colorstereonet.cpp:127-130).
  const TensorDesc& d = d_inputLeft->desc();
  Tensor rgb(d_inputLeft->data(), TensorDesc(1, 3, d.h, d.w));
  Tensor* x = &rgb;
  std::unique_ptr<Tensor> cropped;
  if (m_pwCrop > 0) {  // cuda::cropTensor (stereo/colorstereonet.cpp:71-72), per channel plane
    const int ch = d.h - 2 * m_pwCrop, cw = d.w - 2 * m_pwCrop;
    cropped.reset(new Tensor(TensorDesc(1, 3, ch, cw)));
    for (int c = 0; c < 3; ++c) {
      const float* src = d_inputLeft->data() + ((size_t)c * d.h + m_pwCrop) * d.w + m_pwCrop;
      if (hipMemcpy2DAsync(cropped->data() + (size_t)c * ch * cw, sizeof(float) * cw, src, sizeof(float) * d.w,
                           sizeof(float) * cw, ch, hipMemcpyDeviceToDevice, (hipStream_t)h.stream()) != hipSuccess)
        throw Error(CCS_ERR_HIP, "StereoNet: cropping the pairwise input");
    }
    x = cropped.get();
  }
  x = static_cast<Convolution*>(m_pairwiseOps[0])->forward(x, h);  // + bias + tanh
  Convolution* c2 = static_cast<Convolution*>(m_pairwiseOps[3]);
  Tensor* y = c2->forward(x, h);                                   // + bias + tanh
  const TensorDesc& yd = y->desc();
  Convolution* c3 = static_cast<Convolution*>(m_pairwiseOps[6]);
  check(ccs_pairwise_head(h.get(), y->raw(), prec_code(yd.precision), yd.h, yd.w, 64, c3->getParams()->data(), 2,
                          m_pairwiseOut->data()),
        "StereoNet: pairwise head");
  if (cropped || m_allowGc) h.synchronize();  // the crop (and freed maps) must outlive the launches
  if (m_allowGc) {
    static_cast<Convolution*>(m_pairwiseOps[0])->deleteOutputMemory();
    c2->deleteOutputMemory();
  }
}

Tensor* StereoNet::predictPrepared(Tensor* d_left, Tensor* d_right) { return performPrediction(d_left, d_right); }

Tensor* StereoNet::predictPlanar(const float* d_left, const float* d_right) {
  Handle& h = *m_handle;
  TensorDesc d(1, m_ic, m_ih, m_iw);
  Tensor l(d), r(d);
  check(ccs_pad_normalize(h.get(), d_left, m_ic, m_ih, m_iw, 0, l.data(), nullptr), "StereoNet: normalise");
  check(ccs_pad_normalize(h.get(), d_right, m_ic, m_ih, m_iw, 0, r.data(), nullptr), "StereoNet: normalise");
  Tensor* out = performPrediction(&l, &r);
  h.synchronize();  // l, r are freed on return
  return out;
}

// ----------------------------------------------------------- ColorStereoNet

ColorStereoNet::ColorStereoNet(int numLayers, int in, int ic, int ih, int iw) : StereoNet(numLayers, in, ic, ih, iw) {
  if (ic != 3 && ic != 5) arg_fail("ColorStereoNet: ic must be 3 (RGB) or 5 (RGB + x/y position)");
}

Tensor* ColorStereoNet::predict(const float* d_left_c4, const float* d_right_c4, long pitch) {
  Handle& h = *m_handle;
  if (pitch == 0) pitch = 4L * m_iw;
  TensorDesc d(1, m_ic, m_ih, m_iw);
  Tensor l(d), r(d);
  // C4 -> planar RGB, then per-channel zero mean / unit std (stereo/colorstereonet.cpp:53-68).
  check(ccs_c4_to_planar(h.get(), d_left_c4, m_ih, m_iw, pitch, 3, l.data()), "ColorStereoNet: unpack");
  check(ccs_c4_to_planar(h.get(), d_right_c4, m_ih, m_iw, pitch, 3, r.data()), "ColorStereoNet: unpack");
  check(ccs_pad_normalize(h.get(), l.data(), 3, m_ih, m_iw, 0, l.data(), nullptr), "ColorStereoNet: normalise");
  check(ccs_pad_normalize(h.get(), r.data(), 3, m_ih, m_iw, 0, r.data(), nullptr), "ColorStereoNet: normalise");
  if (m_ic == 5) {  // x/y position planes after the normalised RGB (stereo/colorstereonet.cpp:100-124)
    const size_t plane = (size_t)m_ih * m_iw;
    check(ccs_xy_planes(h.get(), m_ih, m_iw, l.data() + 3 * plane), "ColorStereoNet: x/y planes");
    check(ccs_xy_planes(h.get(), m_ih, m_iw, r.data() + 3 * plane), "ColorStereoNet: x/y planes");
  }
  Tensor* out = performPrediction(&l, &r);
  h.synchronize();
  return out;
}

Tensor* ColorStereoNet::predictHost(const float* left_c4, const float* right_c4) {
  TensorDesc d(1, 4, m_ih, m_iw, Layout::NHWC);
  Tensor l(d), r(d);
  l.copyFromHost(left_c4);
  r.copyFromHost(right_c4);
  return predict(l.data(), r.data());
}

// ------------------------------------------------------------------- Stereo

int label_count(float disp_min, float disp_step, float disp_max, float* disp_max_out) {
  if (disp_step <= 0 || disp_max < disp_min) arg_fail("label_count: need step > 0 and max >= min");
  int depth = (int)((disp_max - disp_min) / disp_step) + 1;
  depth = std::min(256, 1 << (log2i(depth - 1) + 1));
  if (disp_max_out) *disp_max_out = disp_min + (depth - 1) * disp_step;
  return depth;
}

Stereo::Stereo() = default;

Stereo::~Stereo() {
  delete stereoNet_;
  for (Tensor* t : {img1_, img2_, img1_color_, img2_color_, costvolume_, pairwise_, wx_, wy_, out_, labels_, pad1_, pad2_,
                    c1_})
    delete t;
}

void Stereo::initialize(int width, int height, float disp_min, float disp_step, float disp_max) {
  if (width < 1 || height < 1 || disp_step <= 0 || disp_max < disp_min) arg_fail("Stereo::initialize: bad arguments");
  if (!handle_) handle_.reset(new Handle(nullptr));
  width_ = width;
  height_ = height;
  depth_ = label_count(disp_min, disp_step, disp_max, &disp_max_);
  disp_min_ = disp_min;
  disp_step_ = disp_step;
  img1_ = new Tensor(TensorDesc(1, 1, height, width));
  img2_ = new Tensor(TensorDesc(1, 1, height, width));
  img1_color_ = new Tensor(TensorDesc(1, 4, height, width, Layout::NHWC));
  img2_color_ = new Tensor(TensorDesc(1, 4, height, width, Layout::NHWC));
  costvolume_ = new Tensor(TensorDesc(1, depth_, height, width, Layout::NHWC));
  // wx and wy are the two planes of one 1 x 2 x H x W buffer (stereo/stereo.cpp:84-86).
  pairwise_ = new Tensor(TensorDesc(1, 2, height, width));
  wx_ = new Tensor(pairwise_->data(), TensorDesc(1, 1, height, width));
  wy_ = new Tensor(pairwise_->data() + (size_t)width * height, TensorDesc(1, 1, height, width));
  out_ = new Tensor(TensorDesc(1, 1, height, width));
  labels_ = new Tensor(TensorDesc(1, 1, height, width));  // int32 storage
}

void Stereo::setImages(const float* img1, const float* img2) {
  if (!img1_) arg_fail("Stereo: call initialize first");
  img1_->copyFromHost(img1);
  img2_->copyFromHost(img2);
  have_color_ = false;
}

void Stereo::setColorImages(const float* img1_c4, const float* img2_c4) {
  if (!img1_) arg_fail("Stereo: call initialize first");
  img1_color_->copyFromHost(img1_c4);
  img2_color_->copyFromHost(img2_c4);
  // Gray images for the edge weights (stereo/stereo.cpp:120-128).
  Tensor planar(TensorDesc(1, 3, height_, width_));
  for (int i = 0; i < 2; ++i) {
    check(ccs_c4_to_planar(handle_->get(), (i ? img2_color_ : img1_color_)->data(), height_, width_, 4L * width_, 3,
                           planar.data()),
          "Stereo: unpack");
    check(ccs_rgb_to_gray(handle_->get(), planar.data(), height_, width_, (i ? img2_ : img1_)->data()),
          "Stereo: rgb2gray");
  }
  handle_->synchronize();
  have_color_ = true;
}

void Stereo::setWithXy(bool value) {
  if (stereoNet_ && value != with_xy_) arg_fail("Stereo::setWithXy: the network is already built");
  with_xy_ = value;
}

int Stereo::padval() const {
  switch (num_cnn_layers_) {  // stereo/stereo.cpp:204-217
    case 3: return 2;
    case 5: return 3;
    case 7: return 4;
  }
  arg_fail("Stereo: numLayers must be 3, 5 or 7 (stereo/stereo.cpp:198-202)");
}

void Stereo::ensureNet() {
  if (stereoNet_) return;
  const int p = padval();
  const bool pairwise = matching_function_ == COLORCNN_PAIRWISE;
  const bool colour = matching_function_ == COLORCNN || pairwise;
  // Colour models take RGB, plus the x/y position planes with with_xy (stereo/stereo.cpp:237-239).
  const int ic = colour ? (with_xy_ ? 5 : 3) : 1;
  stereoNet_ = colour ? new ColorStereoNet(num_cnn_layers_, 1, ic, height_ + 2 * p, width_ + 2 * p)
                      : new StereoNet(num_cnn_layers_, 1, 1, height_ + 2 * p, width_ + 2 * p);
  stereoNet_->setPrecision(precision_);
  stereoNet_->setVerbose(verbose_);
  // COLORPAIR: the pairwise net writes wx, wy (stereo/stereo.cpp:240-247).
  stereoNet_->initNet(disp_min_, disp_max_, disp_step_, 0, costvolume_, pairwise ? pairwise_ : nullptr);
  std::string params = cnn_parameter_file_base_name_;
  if (params.size() < 4 || params.compare(params.size() - 4, 4, ".npz") != 0) params += ".npz";
  stereoNet_->setAllParams(params);
  stereoNet_->setAllowGc(allowCnnGc_);
}

void Stereo::computeVolume(bool lr) {
  if (!img1_) arg_fail("Stereo: call initialize first");
  if (matching_function_ == LOAD) return;
  const bool colour = matching_function_ == COLORCNN || matching_function_ == COLORCNN_PAIRWISE;
  if (!colour && matching_function_ != GRAYCNN)
    throw Error(CCS_ERR_UNSUPPORTED, "Stereo::computeVolume: matching function not implemented");
  if (colour && !have_color_) arg_fail("Stereo: COLORCNN / COLORPAIR need setColorImages");
  ensureNet();
  Handle& h = stereoNet_->handle();
  const int p = padval();
  const int ic = colour ? 3 : 1;  // image planes; the x/y model adds two position planes
  const int nc = colour && with_xy_ ? 5 : ic;
  const int Hp = height_ + 2 * p, Wp = width_ + 2 * p;
  if (lr)
    stereoNet_->setDisparities(disp_min_, disp_max_, disp_step_, costvolume_);
  else
    stereoNet_->setDisparities(-disp_max_, -disp_min_, disp_step_, costvolume_);
  // Reference image first: left for lr, right for rl (stereo/stereo.cpp:260-318).
  Tensor* src[2];
  std::unique_ptr<Tensor> planar[2];
  for (int i = 0; i < 2; ++i) {
    const bool first_is_left = (i == 0) == lr;
    if (ic == 3) {
      planar[i].reset(new Tensor(TensorDesc(1, 3, height_, width_)));
      check(ccs_c4_to_planar(h.get(), (first_is_left ? img1_color_ : img2_color_)->data(), height_, width_,
                             4L * width_, 3, planar[i]->data()),
            "Stereo: unpack");
      src[i] = planar[i].get();
    } else {
      src[i] = first_is_left ? img1_ : img2_;
    }
  }
  if (!pad1_) {
    pad1_ = new Tensor(TensorDesc(1, nc, Hp, Wp));
    pad2_ = new Tensor(TensorDesc(1, nc, Hp, Wp));
  }
  // Mirror pad + per-channel normalisation in one pass each.
  check(ccs_pad_normalize(h.get(), src[0]->data(), ic, height_, width_, p, pad1_->data(), nullptr), "Stereo: pad");
  check(ccs_pad_normalize(h.get(), src[1]->data(), ic, height_, width_, p, pad2_->data(), nullptr), "Stereo: pad");
  if (nc == 5) {  // position planes of the padded image (stereo/colorstereonet.cpp:100-124)
    const size_t plane = (size_t)Hp * Wp;
    check(ccs_xy_planes(h.get(), Hp, Wp, pad1_->data() + 3 * plane), "Stereo: x/y planes");
    check(ccs_xy_planes(h.get(), Hp, Wp, pad2_->data() + 3 * plane), "Stereo: x/y planes");
  }
  stereoNet_->predictPrepared(pad1_, pad2_);
  h.synchronize();
}

void Stereo::computeWeights(float alpha, float beta, float lambda, bool lr) {
  Handle& h = *handle_;
  check(ccs_edge_weights(h.get(), (lr ? img1_ : img2_)->data(), height_, width_, 0.7f, alpha, beta, lambda,
                         wx_->data(), wy_->data()),
        "Stereo::computeWeights");
}

Tensor* Stereo::getWx(float alpha, float beta, bool left) {
  if (alpha > 0) computeWeights(alpha, beta, 1.f, left);
  handle_->synchronize();
  return wx_;
}

Tensor* Stereo::getWy(float alpha, float beta, bool left) {
  if (alpha > 0) computeWeights(alpha, beta, 1.f, left);
  handle_->synchronize();
  return wy_;
}

void Stereo::computeSlackPropagation(float C1, float C2, float delta, int iterations, float alpha, float beta,
                                     float lambda_reg, bool lr) {
  computeVolume(lr);
  Handle& h = *handle_;
  if (matching_function_ == COLORCNN_PAIRWISE)  // weights from the CNN, times lambda (stereo/stereo.cpp:132-136)
    check(ccs_scale_forward(h.get(), pairwise_->data(), 2L * width_ * height_, lambda_reg), "Stereo: scale weights");
  else
    computeWeights(alpha, beta, lambda_reg, lr);
  Tensor lb(TensorDesc(1, 1, 1, 2));
  if (!c1_) c1_ = new Tensor(TensorDesc(1, depth_, height_, width_, Layout::NHWC));
  check(ccs_slackprop(h.get(), costvolume_->data(), wx_->data(), wy_->data(), height_, width_, depth_, C1, C2,
                      (int)std::lround(delta), iterations, (int*)labels_->raw(), c1_->data(), nullptr,
                      (double*)lb.raw()),
        "Stereo::computeSlackPropagation");
  c1_valid_ = true;
  c1_lr_ = lr;
  // remapSolution (stereo/stereo.cpp:155-158): lr -> min + l step; rl -> -max + l step.
  check(ccs_remap(h.get(), (const int*)labels_->raw(), height_, width_, lr ? disp_min_ : -disp_max_, disp_step_,
                  out_->data()),
        "Stereo: remap");
  h.synchronize();
  if (hipMemcpy(&lower_bound_, lb.raw(), sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    throw Error(CCS_ERR_HIP, "Stereo: reading the bound");
  if (verbose_) std::printf("CRF lower bound: %.6f\n", lower_bound_);
}

void Stereo::fuseQuadFitDirect() {
  if (!c1_valid_) arg_fail("Stereo::fuseQuadFitDirect: call computeSlackPropagation first");
  // The labels of an rl run start at -disp_max (the reference offsets by disp_min in both cases).
  check(ccs_quad_fit_direct(handle_->get(), out_->data(), c1_->data(), height_, width_, depth_,
                            c1_lr_ ? disp_min_ : -disp_max_, disp_step_),
        "Stereo::fuseQuadFitDirect");
  handle_->synchronize();
}

void Stereo::performLeftRightCheck(Tensor* out, Tensor* left, Tensor* right, float th) {
  if (!out || !left || !right) arg_fail("Stereo::performLeftRightCheck: NULL tensor");
  const TensorDesc& d = left->desc();
  if (!(right->desc() == d) || !(out->desc() == d) || d.c != 1 || d.precision != Precision::FP32)
    arg_fail("Stereo::performLeftRightCheck: three H x W float32 maps of one size");
  check(ccs_left_right_check(handle_->get(), left->data(), right->data(), d.h, d.w, th, out->data()),
        "Stereo::performLeftRightCheck");
  handle_->synchronize();
}

void Stereo::computeArgMax(bool lr) {
  computeVolume(lr);
  check(ccs_argmin(handle_->get(), costvolume_->data(), height_, width_, depth_, lr ? disp_min_ : -disp_max_,
                   disp_step_, out_->data()),
        "Stereo::computeArgMax");
  handle_->synchronize();
}

bool Stereo::loadVolumeFromFile(const std::string& filename) {
  // LOAD matching: a .npy cost volume of H x W x D float32/float64 values.
  if (!costvolume_) arg_fail("Stereo: call initialize first");
  HostArray a;
  try {
    a = npz::load_npy(filename);
  } catch (const Error&) {
    return false;
  }
  if (a.size() != (size_t)height_ * width_ * depth_) return false;
  costvolume_->copyFromHost(a.data.data());
  return true;
}

}  // namespace ccs
