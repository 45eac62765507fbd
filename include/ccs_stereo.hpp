// C++ operator / network API of the MI355X stereo library, over the C ABI of
// ccs.h.  It keeps the reference's class names, constructor parameters and
// method sets so that code written against the reference's operator layer
// switches by changing types only:
//
//   reference                              here
//   iu::TensorGpu_32f                      ccs::Tensor (n, c, h, w, layout, device memory)
//   cudnnTensorDescriptor_t                ccs::TensorDesc
//   cudnnHandle_t                          ccs::Handle (one HIP stream)
//   Operator, Convolution, Bias,           same names (stereo/operators/*.h)
//   Activation, Softmax, StereoCorrelation,
//   SlackProp
//   StereoNet, ColorStereoNet              same names (stereo/stereonet.h, colorstereonet.h)
//   Stereo                                 same name (stereo/stereo.h)
//
// Errors: the reference prints and calls exit(-1); every failure here throws
// ccs::Error carrying the C-ABI status and message.
//
// Layout: convolution outputs and correlation inputs are channel-last
// (NHWC) with the channel stride rounded up to 8 -- the layout the MFMA kernels
// read.  Tensor::toNCHW() produces the reference's planar layout when a caller
// needs it.  Cost volumes are NHWC (H x W x D) as in the reference.
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ccs.h"

namespace ccs {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Throws ccs::Error(rc, "<what>: <ccs_last_error()>") unless rc == CCS_OK.
void check(int rc, const char* what);

enum class Layout { NCHW, NHWC };
enum class Precision { FP32, BF16 };

// Shape + memory layout (replaces cudnnTensorDescriptor_t).  NHWC tensors may
// carry a channel stride cs > c; BF16 is used for channel-last feature maps only.
struct TensorDesc {
  int n = 1, c = 1, h = 1, w = 1;
  Layout layout = Layout::NCHW;
  int cs = 0;  // channel stride of NHWC tensors (0 = c)
  Precision precision = Precision::FP32;

  TensorDesc() = default;
  TensorDesc(int n_, int c_, int h_, int w_, Layout l = Layout::NCHW, int cs_ = 0,
             Precision p = Precision::FP32)
      : n(n_), c(c_), h(h_), w(w_), layout(l), cs(cs_), precision(p) {}
  int cstride() const { return cs > 0 ? cs : c; }
  size_t elements() const;  // stored elements, pad channels included
  size_t bytes() const;
  bool operator==(const TensorDesc& o) const;
};

// One ccs handle bound to one HIP stream (replaces cudnnHandle_t).  The handle
// owns the library's grow-only scratch; use one per host thread.
class Handle {
 public:
  explicit Handle(void* hip_stream = nullptr);
  ~Handle();
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  ccs_handle_t get() const { return h_; }
  void* stream() const { return stream_; }
  void setStream(void* hip_stream);
  void synchronize();

 private:
  ccs_handle_t h_ = nullptr;
  void* stream_ = nullptr;
};

// Device tensor (replaces iu::TensorGpu_32f).  Owns its memory unless it wraps
// an external pointer.
class Tensor {
 public:
  explicit Tensor(const TensorDesc& d);
  Tensor(int n, int c, int h, int w, Layout layout = Layout::NCHW);
  Tensor(void* external, const TensorDesc& d);  // not owned
  ~Tensor();
  Tensor(const Tensor&) = delete;
  Tensor& operator=(const Tensor&) = delete;

  float* data() { return static_cast<float*>(ptr_); }
  const float* data() const { return static_cast<const float*>(ptr_); }
  void* raw() { return ptr_; }
  const void* raw() const { return ptr_; }
  const TensorDesc& desc() const { return desc_; }
  int samples() const { return desc_.n; }
  int channels() const { return desc_.c; }
  int height() const { return desc_.h; }
  int width() const { return desc_.w; }
  Layout memoryLayout() const { return desc_.layout; }
  bool ownsMemory() const { return owned_; }

  // Synchronous host <-> device copies of the stored elements (desc().elements()).
  void copyFromHost(const void* host);
  void copyToHost(void* host) const;
  // The tensor as float32 planar N x C x H x W on the host (converts layout and bf16).
  std::vector<float> toNCHW(Handle& handle) const;

 private:
  TensorDesc desc_;
  void* ptr_ = nullptr;
  bool owned_ = false;
};

// A float32 array read from a .npz archive (npy v1/v2/v3, '<f4', C order).
struct HostArray {
  std::vector<int> shape;
  std::vector<float> data;
  size_t size() const;
};

// NumPy .npz archives as the reference's parameter files store them (ZIP,
// stored or deflated members arr_0.npy, arr_1.npy, ...).  Replaces cnpy.
namespace npz {
std::map<std::string, HostArray> load(const std::string& path);
HostArray load_array(const std::string& path, const std::string& key);
HostArray load_npy(const std::string& path);  // one .npy file
void save_npy(const std::string& path, const HostArray& a);
void save(const std::string& path, const std::vector<std::pair<std::string, HostArray>>& arrays);
}  // namespace npz

// ---------------------------------------------------------------------------
// Operators (stereo/operators/operator.h:26-56)

class Operator {
 public:
  enum Type { CONVOLUTION, BIAS, ACTIVATION, SOFTMAX, STEREO_CORRELATION, SLACKPROP };

  Operator(const TensorDesc& inTensorDesc, Type type);
  Operator(int in, int ic, int ih, int iw, Type type);
  Operator(const std::vector<TensorDesc>& inTensorDesc, Type type);
  virtual ~Operator();

  virtual Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) = 0;
  Tensor* forward(Tensor* d_input, Handle& handle);
  virtual TensorDesc outTensorDesc() { return m_inTensorDesc[0]; }
  virtual Tensor* getParams() { return nullptr; }
  // Loads `array` of the .npz at npzPath into getParams(): the element count
  // must match, and a 4-D array must match the parameter's shape exactly
  // (stereo/operators/operator.cpp:55-85).
  void setParams(const std::string& npzPath, const std::string& array);
  void setParams(const HostArray& h_params);
  // The checks of setParams for a parameter of shape `want` (host only):
  // throws ccs::Error when h_params does not fit.
  static void checkParams(const HostArray& h_params, const TensorDesc& want);
  Type type() const { return m_type; }

 protected:
  virtual void paramsChanged() {}
  std::vector<TensorDesc> m_inTensorDesc;

 private:
  Type m_type;
};

// stereo/operators/convolution.h:25-66.  CONVOLUTION flips the filter (true
// convolution, the reference's CUDNN_CONVOLUTION); CROSS_CORRELATION does not.
// pad 0, stride 1.  Input: N=1, a planar float32 image (C <= 8) or a
// channel-last feature map; output: channel-last, stride round_up(K, 8), in the
// operator's precision (FP32 unless setPrecision(BF16)).
class Convolution : public Operator {
 public:
  enum Mode { CONVOLUTION, CROSS_CORRELATION };

  Convolution(Mode convMode, const TensorDesc& inTensorDesc, int fk, int fc, int fh, int fw,
              Tensor* d_out = nullptr, int pad_x = 0, int pad_y = 0);
  Convolution(Mode convMode, int in, int ic, int ih, int iw, int fk, int fc, int fh, int fw,
              Tensor* d_out = nullptr, int pad_x = 0, int pad_y = 0);
  ~Convolution() override;

  Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) override;
  using Operator::forward;
  Tensor* getOutput() { return m_output; }
  TensorDesc outTensorDesc() override;
  Tensor* getParams() override { return m_filter; }
  void deleteOutputMemory();

  // MI355X additions: arithmetic of the MFMA path, and the bias + tanh epilogue
  // that StereoNet fuses into this convolution (the separate Bias and
  // Activation operators are then skipped by the network).
  void setPrecision(Precision p);
  Precision precision() const { return m_precision; }
  void setFusedEpilogue(const Tensor* bias, bool tanh);

 private:
  void initialize(Tensor* d_out, int pad_x, int pad_y);
  void allocateOutputMemory(Tensor* d_out);
  void paramsChanged() override { m_packedValid = false; }

  Mode m_convMode;
  int m_fk, m_fc, m_fh, m_fw;
  Precision m_precision = Precision::FP32;
  Tensor* m_filter = nullptr;
  Tensor* m_packed = nullptr;
  bool m_packedValid = false;
  Precision m_packedPrecision = Precision::FP32;
  const Tensor* m_fusedBias = nullptr;
  bool m_fusedTanh = false;
  bool m_extOutPtr = false;
  Tensor* m_output = nullptr;
};

// stereo/operators/bias.h: y += b per channel, in place (returns its input).
class Bias : public Operator {
 public:
  Bias(const TensorDesc& inTensorDesc, const TensorDesc& biasTensorDesc);
  Bias(const TensorDesc& inTensorDesc, int fk, int fc, int fh, int fw);
  ~Bias() override;
  Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) override;
  using Operator::forward;
  Tensor* getParams() override { return m_bias; }

 private:
  Tensor* m_bias = nullptr;
};

// stereo/operators/activation.h: tanh in place (returns its input).
class Activation : public Operator {
 public:
  enum Mode { TANH };
  Activation(const TensorDesc& inOutTensorDesc, Mode activationMode = TANH, double reluCeiling = 0.);
  Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) override;
  using Operator::forward;
};

// stereo/operators/softmax.h: softmax over the channels of an NHWC volume, in
// place (CHANNEL mode, ACCURATE).  setNegate(true) folds the network's
// cost = -p into the same pass.
class Softmax : public Operator {
 public:
  enum Mode { CHANNEL };
  Softmax(const TensorDesc& inOutTensorDesc, Mode mode = CHANNEL);
  Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) override;
  using Operator::forward;
  void setNegate(bool v) { m_negate = v; }

 private:
  bool m_negate = false;
};

// stereo/operators/stereocorrelation.h: the correlation cost volume over the
// offsets -(min_disp + k step); output N=1, H x W x D NHWC float32.  Inputs are
// the two towers' channel-last feature maps.  setNegSoftmaxEpilogue(true) fuses
// Softmax + negate into the same kernel (the network does this).
class StereoCorrelation : public Operator {
 public:
  StereoCorrelation(const std::vector<TensorDesc>& inTensorDesc, Tensor* d_out = nullptr,
                    float min_disp = 0, float max_disp = 0, float step = 0, int rect_corr = 0);
  ~StereoCorrelation() override;
  Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) override;
  using Operator::forward;
  TensorDesc outTensorDesc() override;
  void setDisparities(float min_disp, float max_disp, float step, Tensor* d_out = nullptr);
  void setRectificationCorrection(int value) { m_rectCorr = value; }
  void setNegSoftmaxEpilogue(bool v) { m_negSoftmax = v; }
  int numDisparities() const { return m_numDisps; }

  // The reference's raw-pointer statics (stereocorrelation.h:38-44).
  static void forward(float* input0, float* input1, float* disparities, float* output, int in, int ic,
                      int ih, int iw, int lenDisps, Layout memoryLayout, int rectCorr,
                      Handle* handle = nullptr);
  static void backward(float* d_outGrad0, float* d_outGrad1, float* d_inGrad, float* d_im0, float* d_im1,
                       float* d_disparities, int n, int c, int h, int w, int numDisps, Layout memoryLayout,
                       Handle* handle = nullptr);

 private:
  void allocateOutput(Tensor* d_out);
  int m_rectCorr = 0;
  bool m_negSoftmax = false;
  int m_dispMin = 0, m_dispStep = 1, m_numDisps = 0;
  bool m_extOutPtr = false;
  Tensor* m_output = nullptr;
};

// stereo/operators/slackprop.h: SlackProp CRF inference.  forward(inputs) takes
// {cost volume (1, D, H, W as NHWC), wx (H x W), wy (H x W)} and returns the
// labels as float (1 x 1 x H x W), like the reference's static forward.
class SlackProp : public Operator {
 public:
  explicit SlackProp(const TensorDesc& inTensorDesc);
  ~SlackProp() override;
  Tensor* forward(std::vector<Tensor*>& d_inputs, Handle& handle) override;
  using Operator::forward;
  TensorDesc outTensorDesc() override;
  void setLambda(float value) { m_lambda = value; }
  void setP1(float value) { m_P1 = value; }
  void setP2(float value) { m_P2 = value; }
  void setDelta(int value) { m_delta = value; }
  void setIterations(int value) { m_iters = value; }
  float lowerBound() const { return m_lowerBound; }

  // The per-edge ("full") interface (slackprop.h:40): costVolume 1 x D x H x W
  // (NHWC), pwVolume 1 x 2 levels x H x W (NCHW planes P1x, P1y, P2x, P2y, or
  // P x, P y for the Potts form), output 1 x 1 x H x W labels as float.
  static float forward(Tensor* costVolume, Tensor* pwVolume, Tensor* output, int max_iter = 5,
                       Handle* handle = nullptr);
  // The reference's raw-pointer statics (slackprop.h:43-47).
  static float forward(float* costVolume, float* wx, float* wy, float* output, int in, int ic, int ih, int iw,
                       float P1, float P2, int max_iter = 5, Handle* handle = nullptr);
  static float forwardVolumeOut(float* costVolume, float* wx, float* wy, float* output, int in, int ic,
                                int ih, int iw, float P1, float P2, int max_iter, int delta,
                                Handle* handle = nullptr);

 private:
  float m_P1 = 1.0f, m_P2 = 7.0f, m_lambda = 0.18f, m_lowerBound = -1.0f;
  int m_delta = 2, m_iters = 5;
  Tensor* m_output = nullptr;
};

// ---------------------------------------------------------------------------
// Networks (stereo/stereonet.h, stereo/colorstereonet.h)

class StereoNet {
 public:
  // numLayers in {3, 5, 7}: a 3x3 layer then 2x2 layers, 100 channels
  // (stereo/stereonet.cpp:65-141).  in must be 1; (ih, iw) is the padded size.
  StereoNet(int numLayers, int in, int ic, int ih, int iw);
  virtual ~StereoNet();
  StereoNet(const StereoNet&) = delete;
  StereoNet& operator=(const StereoNet&) = delete;

  void setDisparities(float min_disp, float max_disp, float step, Tensor* d_out = nullptr);
  void setAllParams(const std::string& npzPath);
  // Host-only forms of the parameter walk.  parameterShapes lists, in the order
  // of the arr_<i> keys, the parameters setAllParams loads for this
  // architecture: per layer the filter (K x C x KH x KW) and the bias (1 x K x
  // 1 x 1), the tower's layers first, then the pairwise network's (whose 1x1
  // head has no bias).  checkParameterFile walks an npz file against it with
  // setParams's checks, touching no device; it returns the keys of the arrays
  // left unread (the trailing learned L1, L2 of CRF models) and throws
  // ccs::Error where setAllParams would.
  static std::vector<TensorDesc> parameterShapes(int numLayers, int ic, bool pairwise);
  static std::vector<std::string> checkParameterFile(const std::string& npzPath, int numLayers, int ic,
                                                     bool pairwise);
  void setVerbose(bool val) { m_verbose = val; }
  void setAllowGc(bool value) { m_allowGc = value; }
  // With d_pairwiseOut (a 1 x 2 x H x W float32 planar tensor, H x W the
  // towers' output size) the network also builds the pairwise CNN of the
  // COLORPAIR models (colour input only): performPrediction then writes wx, wy.
  virtual void initNet(float min_disp = 0, float max_disp = 0, float step = 0, int rect_corr = 0,
                       Tensor* d_unaryOut = nullptr, Tensor* d_pairwiseOut = nullptr);
  // MI355X additions: tower arithmetic and the handle the network runs on.
  // FP32 by default, as in the reference; Precision::BF16 (MFMA, FP32
  // accumulation) is opt-in, about 3x faster with cost-volume errors near 1e-3.
  void setPrecision(Precision p);
  Handle& handle() { return *m_handle; }
  const std::vector<Operator*>& leftOps() const { return m_leftOps; }
  const std::vector<Operator*>& pairwiseOps() const { return m_pairwiseOps; }

  // Gray (ic = 1) or planar prediction from padded, un-normalised device images
  // (C x ih x iw float32): normalises each image, runs both towers and the
  // correlation with the fused -softmax.  Returns the H x W x D cost volume.
  Tensor* predictPlanar(const float* d_left, const float* d_right);
  // Both towers + correlation on images already padded and normalised
  // (1 x ic x ih x iw float32 planar tensors).
  Tensor* predictPrepared(Tensor* d_left, Tensor* d_right);

 protected:
  Tensor* performPrediction(Tensor* d_inputLeft, Tensor* d_inputRight);
  void initPairwise(Tensor* d_pairwiseOut);
  void runPairwise(Tensor* d_inputLeft);

  std::unique_ptr<Handle> m_handle;
  // The right tower runs on a second stream beside the left one (HIP events
  // order it after the inputs and the correlation after it).
  std::unique_ptr<Handle> m_handle2;
  void* m_stream2 = nullptr;
  void* m_evIn = nullptr;
  void* m_evOut = nullptr;
  std::vector<Operator*> m_leftOps, m_rightOps, m_lrOps, m_pairwiseOps;
  Tensor* m_pairwiseOut = nullptr;  // 1 x 2 x H x W: wx, wy (caller-owned)
  int m_pwCrop = 0;                 // pixels cropped from each side of the pairwise net's input
  int m_numLayers, m_in, m_ic, m_ih, m_iw;
  bool m_verbose = false, m_initialized = false, m_allowGc = false;
  Precision m_precision = Precision::FP32;
};

class ColorStereoNet : public StereoNet {
 public:
  ColorStereoNet(int numLayers, int in, int ic, int ih, int iw);
  // Padded RGBA images (ih x iw float4 texels, row pitch in floats; 0 = 4 iw),
  // on the device or on the host.  Returns the H x W x D cost volume.
  Tensor* predict(const float* d_left_c4, const float* d_right_c4, long pitch = 0);
  Tensor* predictHost(const float* left_c4, const float* right_c4);
};

// ---------------------------------------------------------------------------
// Facade (stereo/stereo.h)

enum MatchingFunction { COLORCNN, GRAYCNN, CONCATCNN, GRAYCNN_PAIRWISE, COLORCNN_PAIRWISE,
                        CONCATCNN_PAIRWISE, CENSUS, LOAD };

// Number of labels of the reference's facade for a disparity range: n = (max -
// min)/step + 1 rounded up to a power of two, at most 256; *disp_max_out gets
// min + (labels - 1) step (stereo/stereo.cpp:64-73).
int label_count(float disp_min, float disp_step, float disp_max, float* disp_max_out);

class Stereo {
 public:
  Stereo();
  virtual ~Stereo();
  Stereo(const Stereo&) = delete;
  Stereo& operator=(const Stereo&) = delete;

  // depth = min(256, 2^ceil(log2(n))) for n = (max - min)/step + 1 labels;
  // disp_max is then min + (depth - 1) step (stereo/stereo.cpp:64-73).
  void initialize(int width, int height, float disp_min, float disp_step, float disp_max);
  // Host images, row-major: gray H x W, color H x W RGBA.
  void setImages(const float* img1, const float* img2);
  void setColorImages(const float* img1_c4, const float* img2_c4);
  void computeSlackPropagation(float C1, float C2, float delta, int iterations, float alpha, float beta,
                               float lambda_reg, bool lr);
  void computeArgMax(bool lr);
  // QuadDirect sub-pixel refinement of the CRF output (stereo/stereo.cpp:406-409):
  // a parabola through the three values around each label of the final row
  // unaries of the last computeSlackPropagation (the solver's get_C1 here).
  void fuseQuadFitDirect();
  // stereo/stereo.cpp:448-455: out = left where |left + right(x - left)| <= th,
  // else -1 (right: the right-to-left disparities, negative).
  void performLeftRightCheck(Tensor* out, Tensor* left, Tensor* right, float th);
  Tensor* getC1() { return c1_; }  // H x W x D final row unaries of the last CRF run
  void computeVolume(bool lr);
  void computeWeights(float alpha, float beta, float lambda, bool lr);
  // stereo/stereo.cpp:425-437: with alpha > 0 the edge weights are recomputed
  // with lambda = 1 first; otherwise the current wx / wy are returned.
  Tensor* getWx(float alpha, float beta, bool left);
  Tensor* getWy(float alpha, float beta, bool left);
  void setMatchingFunction(MatchingFunction function) { matching_function_ = function; }
  void setNumLayers(int num_layers) { num_cnn_layers_ = num_layers; }
  void setParameterFile(const std::string& base_name) { cnn_parameter_file_base_name_ = base_name; }
  bool loadVolumeFromFile(const std::string& filename);  // .npy, H x W x D float32
  void setVerbose(bool val) { verbose_ = val; }
  void setAllowCnnGc(bool value) { allowCnnGc_ = value; }
  void setWithXy(bool value);
  void setPrecision(Precision p) { precision_ = p; }

  Tensor* getOutput() { return out_; }        // H x W disparities, float32
  Tensor* getVolume() { return costvolume_; }  // H x W x D, float32
  Tensor* getWx() { return wx_; }
  Tensor* getWy() { return wy_; }
  Tensor* getLabels() { return labels_; }  // H x W int32
  int getDepth() const { return depth_; }
  int width() const { return width_; }
  int height() const { return height_; }
  float dispMin() const { return disp_min_; }
  float dispMax() const { return disp_max_; }
  float dispStep() const { return disp_step_; }
  double lowerBound() const { return lower_bound_; }
  Handle& handle() { return *handle_; }

 private:
  int padval() const;
  void ensureNet();

  std::unique_ptr<Handle> handle_;
  bool verbose_ = false, allowCnnGc_ = false, have_color_ = false, with_xy_ = false;
  int width_ = 0, height_ = 0, depth_ = 0;
  float disp_min_ = 0, disp_step_ = 1, disp_max_ = 0;
  int num_cnn_layers_ = 3;
  std::string cnn_parameter_file_base_name_ = "params";
  MatchingFunction matching_function_ = COLORCNN;
  Precision precision_ = Precision::FP32;
  double lower_bound_ = 0.0;
  Tensor *img1_ = nullptr, *img2_ = nullptr;              // gray, H x W
  Tensor *img1_color_ = nullptr, *img2_color_ = nullptr;  // RGBA, H x W x 4
  Tensor* costvolume_ = nullptr;
  Tensor *pairwise_ = nullptr, *wx_ = nullptr, *wy_ = nullptr;
  Tensor *out_ = nullptr, *labels_ = nullptr;
  Tensor* c1_ = nullptr;  // final row unaries of the last CRF run (allocated on first use)
  bool c1_valid_ = false, c1_lr_ = true;
  Tensor *pad1_ = nullptr, *pad2_ = nullptr;
  StereoNet* stereoNet_ = nullptr;
};

// ---------------------------------------------------------------------------
// Configuration (stereo/stereo_img.cpp:47-103): "key = value" lines, '#'
// comments; options given on the command line take precedence over the file.

class Options {
 public:
  Options();  // the reference's defaults
  // Parses argv (--key value, --key=value, -o value, -g value, positional im0 im1),
  // then the file named by --config-file for keys not given on the command line.
  // Throws ccs::Error on unknown keys or malformed values.
  void parse(int argc, const char* const* argv);
  void parseFile(const std::string& path, bool override_existing = false);
  void set(const std::string& key, const std::string& value, bool from_cli = true);
  bool has(const std::string& key) const;
  std::string str(const std::string& key) const;
  double num(const std::string& key) const;
  int integer(const std::string& key) const;
  bool flag(const std::string& key) const;
  static std::string usage();

 private:
  std::map<std::string, std::string> values_;
  std::map<std::string, bool> explicit_;
};

// Image files: PFM (float), PGM/PPM (8/16-bit binary) and PNG (8/16-bit gray,
// RGB, RGBA; zlib).  Images load as RGBA float in [0, 1]; saving writes a gray
// float map.
struct Image {
  int width = 0, height = 0;
  std::vector<float> rgba;  // H x W x 4
};
Image imread_c4(const std::string& path);
void save_pfm(const std::string& path, const float* data, int width, int height);
std::vector<float> load_pfm(const std::string& path, int* width, int* height);
// 16-bit PNG of disparity * 256 (KITTI convention), or 8-bit scaled to the max when !kitti.
void save_png_disparity(const std::string& path, const float* data, int width, int height, bool kitti = true);

}  // namespace ccs
