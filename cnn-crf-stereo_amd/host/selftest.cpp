// ccs_selftest: host-only entry points of the C++ API for the CPU test suite
// (no device work): .npz / .npy / PNG / PFM I/O, option parsing and the label
// count rule.  Each subcommand prints plain text that the tests compare.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ccs_stereo.hpp"

using namespace ccs;

static void dump(const std::string& key, const HostArray& a) {
  double sum = 0, sq = 0;
  for (float v : a.data) {
    sum += v;
    sq += (double)v * v;
  }
  std::printf("%s shape", key.c_str());
  for (int d : a.shape) std::printf(" %d", d);
  std::printf(" n %zu sum %.9g sq %.9g first %.9g last %.9g\n", a.data.size(), sum, sq,
              a.data.empty() ? 0.0 : a.data.front(), a.data.empty() ? 0.0 : a.data.back());
}

static int run(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: ccs_selftest npz-dump|npy-dump|npz-write|npy-write|options|config|params|labels|imread|pfm-write|png-write ...\n");
    return 1;
  }
  const std::string cmd = argv[1];
  if (cmd == "npz-dump" && argc == 3) {
    for (const auto& [k, a] : npz::load(argv[2])) dump(k, a);
  } else if (cmd == "npy-dump" && argc == 3) {
    dump("npy", npz::load_npy(argv[2]));
  } else if ((cmd == "npz-write" || cmd == "npy-write") && argc == 3) {
    // arr_0: 2 x 3 x 2 x 2 ramp, arr_1: 5-vector, arr_2: a 0-d array.
    HostArray a, b, c;
    a.shape = {2, 3, 2, 2};
    for (int i = 0; i < 24; ++i) a.data.push_back(0.5f * i - 3.f);
    b.shape = {5};
    b.data = {1.f, -2.f, 3.5f, 1e-3f, 7.f};
    c.data = {42.f};
    if (cmd == "npz-write") npz::save(argv[2], {{"arr_0", a}, {"arr_1", b}, {"arr_2", c}});
    else npz::save_npy(argv[2], a);
  } else if (cmd == "options") {
    Options o;
    o.parse(argc - 1, argv + 1);
    for (const char* k : {"im0", "im1", "disp-min", "disp-max", "disp-step", "matching", "inference", "lambda",
                          "L1", "L2", "delta", "alpha", "beta", "crf-iterations", "num-layers", "output-file",
                          "gpu", "parameter-file", "precision"})
      std::printf("%s=%s\n", k, o.has(k) ? o.str(k).c_str() : "<unset>");
  } else if (cmd == "config" && argc == 3) {
    // A model .cfg through Options::parseFile, as stereo_img --config-file reads it.
    Options o;
    o.parseFile(argv[2]);
    for (const char* k : {"matching", "num-layers", "inference", "lambda", "L1", "L2", "delta", "alpha", "beta",
                          "disp-min", "disp-step", "disp-max", "crf-iterations", "with-xy"})
      std::printf("%s=%s\n", k, o.has(k) ? o.str(k).c_str() : "<unset>");
  } else if (cmd == "params" && argc == 6) {
    // StereoNet's parameter walk over an npz file, host only: numLayers, input
    // channels, pairwise network (0/1).
    const auto unread = StereoNet::checkParameterFile(argv[2], atoi(argv[3]), atoi(argv[4]), atoi(argv[5]) != 0);
    std::printf("ok %zu", StereoNet::parameterShapes(atoi(argv[3]), atoi(argv[4]), atoi(argv[5]) != 0).size());
    for (const auto& k : unread) std::printf(" %s", k.c_str());
    std::printf("\n");
  } else if (cmd == "labels" && argc == 5) {
    float mx = 0;
    const int n = label_count((float)atof(argv[2]), (float)atof(argv[3]), (float)atof(argv[4]), &mx);
    std::printf("%d %.6g\n", n, mx);
  } else if (cmd == "imread" && argc == 3) {
    const Image im = imread_c4(argv[2]);
    std::printf("%d %d\n", im.width, im.height);
    for (size_t i = 0; i < im.rgba.size(); ++i) std::printf("%.7g%c", im.rgba[i], (i + 1) % 4 ? ' ' : '\n');
  } else if ((cmd == "pfm-write" || cmd == "png-write") && argc == 3) {
    // 5 x 3 map: value = 0.25 * (x + 10 y) - 1 (negative entries clamp to 0 in the PNG).
    std::vector<float> d(15);
    for (int y = 0; y < 3; ++y)
      for (int x = 0; x < 5; ++x) d[y * 5 + x] = 0.25f * (x + 10 * y) - 1.f;
    if (cmd == "pfm-write") save_pfm(argv[2], d.data(), 5, 3);
    else save_png_disparity(argv[2], d.data(), 5, 3, true);
  } else {
    std::fprintf(stderr, "ccs_selftest: bad command\n");
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const Error& e) {
    std::fprintf(stderr, "ccs_selftest: %s\n", e.what());
    return 2;
  }
}
