// stereo_img: the command-line front end of stereo/stereo_img.cpp over the
// MI355X library.  Same options and config-file handling; outputs
// <output-file>.png (16-bit disparity * 256), .pfm, _wx.pfm, _wy.pfm for CRF
// inference, .png/.pfm for ARGMAX, and the .npy volume plus weights for NONE.
// Matching: COLORCNN, GRAYCNN, COLORPAIR (pairwise CNN weights), LOAD.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ccs_stereo.hpp"

using namespace ccs;

static std::vector<float> to_host(Tensor* t) {
  std::vector<float> v(t->desc().elements());
  t->copyToHost(v.data());
  return v;
}

static int run(int argc, char** argv) {
  Options o;
  o.parse(argc, argv);
  if (o.has("help") || !o.has("im0") || !o.has("im1")) {
    std::fputs(Options::usage().c_str(), stdout);
    return 1;
  }
  if (hipSetDevice(o.integer("gpu")) != hipSuccess) throw Error(CCS_ERR_HIP, "hipSetDevice failed");
  const std::string out = o.str("output-file");
  const Image I1 = imread_c4(o.str("im0")), I2 = imread_c4(o.str("im1"));
  if (I1.width != I2.width || I1.height != I2.height) throw Error(CCS_ERR_ARG, "the two images differ in size");

  Stereo stereo;
  stereo.setAllowCnnGc(o.flag("cnn-gc"));
  stereo.setWithXy(o.flag("with-xy"));
  stereo.initialize(I1.width, I1.height, (float)o.num("disp-min"), (float)o.num("disp-step"),
                    (float)o.num("disp-max"));
  stereo.setVerbose(false);
  const std::string prec = o.str("precision");
  if (prec != "bf16" && prec != "fp32") throw Error(CCS_ERR_ARG, "precision must be bf16 or fp32");
  stereo.setPrecision(prec == "bf16" ? Precision::BF16 : Precision::FP32);

  const std::string matching = o.str("matching");
  float alpha = (float)o.num("alpha");
  if (matching == "COLORCNN" || matching == "GRAYCNN" || matching == "COLORPAIR") {
    stereo.setNumLayers(o.integer("num-layers"));
    stereo.setParameterFile(o.str("parameter-file"));
    stereo.setMatchingFunction(matching == "COLORCNN" ? COLORCNN : matching == "GRAYCNN" ? GRAYCNN : COLORCNN_PAIRWISE);
    if (matching == "COLORPAIR") alpha = 0.f;  // the weights come from the pairwise CNN (stereo_img.cpp:151-156)
  } else if (matching == "LOAD") {
    stereo.setMatchingFunction(LOAD);
    if (!stereo.loadVolumeFromFile(o.str("volume")))
      throw Error(CCS_ERR_IO, "cannot load the volume " + o.str("volume"));
  } else if (matching == "CENSUS") {
    throw Error(CCS_ERR_UNSUPPORTED, "matching " + matching + " is not implemented by this build");
  } else {
    throw Error(CCS_ERR_ARG, "Unknown matching function " + matching);
  }
  stereo.setColorImages(I1.rgba.data(), I2.rgba.data());

  const float beta = (float)o.num("beta"), lambda = (float)o.num("lambda");
  const std::string inference = o.str("inference");
  const auto t0 = std::chrono::steady_clock::now();
  if (inference == "CRF") {
    const std::string refinement = o.str("refinement");
    if (refinement != "NONE" && refinement != "QuadDirect")
      throw Error(CCS_ERR_UNSUPPORTED, "refinement " + refinement);
    stereo.computeSlackPropagation((float)o.num("L1"), (float)o.num("L2"), (float)o.num("delta"),
                                   o.integer("crf-iterations"), alpha, beta, lambda, true);
    if (refinement == "QuadDirect") stereo.fuseQuadFitDirect();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // As the reference saves them: recomputed with lambda = 1 when alpha > 0 (stereo.cpp:425-437).
    std::vector<float> d = to_host(stereo.getOutput()), wx = to_host(stereo.getWx(alpha, beta, true)),
                       wy = to_host(stereo.getWy(alpha, beta, true));
    save_png_disparity(out + ".png", d.data(), I1.width, I1.height, true);
    save_pfm(out + ".pfm", d.data(), I1.width, I1.height);
    save_pfm(out + "_wx.pfm", wx.data(), I1.width, I1.height);
    save_pfm(out + "_wy.pfm", wy.data(), I1.width, I1.height);
    std::printf("stereo_img: %dx%d D=%d CRF lower bound %.6f, %.2f ms (includes first-call setup)\n", I1.width,
                I1.height, stereo.getDepth(), stereo.lowerBound(), ms);
  } else if (inference == "ARGMAX") {
    stereo.computeArgMax(true);
    std::vector<float> d = to_host(stereo.getOutput());
    save_png_disparity(out + ".png", d.data(), I1.width, I1.height, true);
    save_pfm(out + ".pfm", d.data(), I1.width, I1.height);
  } else if (inference == "NONE") {
    stereo.computeVolume(true);
    HostArray wx, wy, vol;
    wx.shape = wy.shape = {I1.height, I1.width};
    wx.data = to_host(stereo.getWx(alpha, beta, true));
    wy.data = to_host(stereo.getWy(alpha, beta, true));
    vol.shape = {I1.height, I1.width, stereo.getDepth()};
    vol.data = to_host(stereo.getVolume());
    npz::save_npy(out + "_wx.npy", wx);
    npz::save_npy(out + "_wy.npy", wy);
    npz::save_npy(out + ".npy", vol);
  } else {
    throw Error(CCS_ERR_ARG, "Unknown inference method " + inference);
  }
  return 0;
}

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const Error& e) {
    std::fprintf(stderr, "stereo_img: %s\n", e.what());
    return 2;
  }
}
