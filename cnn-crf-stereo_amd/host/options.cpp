// Command-line / config-file options of stereo_img (stereo/stereo_img.cpp:47-103):
// the same keys and defaults; a value given on the command line wins over the
// config file, as with boost::program_options store order in the reference.
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "../../include/ccs_stereo.hpp"

namespace ccs {
namespace {

struct Spec {
  const char* key;
  const char* def;  // nullptr: no default
  const char* help;
};

const Spec kSpecs[] = {
    {"help", nullptr, "produce help message"},
    {"im0", nullptr, "Left Image"},
    {"im1", nullptr, "Right Image"},
    {"disp-min", "1", "Minimum disparity"},
    {"disp-step", "1", "Disparity step"},
    {"disp-max", "128", "Maximum disparity"},
    {"matching", "CENSUS", "Matching method: COLORCNN, GRAYCNN, COLORPAIR, LOAD (CENSUS: not implemented)"},
    {"volume", "", "Volume file (.npy, H x W x D) when matching is LOAD"},
    {"inference", "CRF", "Inference method: CRF, ARGMAX, NONE"},
    {"refinement", "NONE", "Continuous refinement: NONE, QuadDirect (THuberQuad: not implemented)"},
    {"alpha", "20.0", "Parameter for edge image calculation"},
    {"beta", "0.8", "Parameter for edge image calculation"},
    {"gpu", "0", "GPU to use"},
    {"output-file", "./output", "Base name (without extension) of the output files"},
    {"config-file", nullptr, "Provide all options in a config file"},
    {"num-layers", "7", "Number of layers for CNN matching (3, 5, 7)"},
    {"parameter-file", "params", "Parameters for CNN matching (.npz, extension optional)"},
    {"with-xy", "false", "Use x- and y-coordinates (5-channel colour models)"},
    {"filter-size", "7", "Filter size for Census matching"},
    {"lambda", "0.4", "Weighting of regularization term"},
    {"L1", "1", "Parameter for CRF inference"},
    {"L2", "12", "Parameter for CRF inference"},
    {"delta", "2", "Parameter for CRF inference"},
    {"crf-iterations", "4", "Parameter for CRF inference"},
    {"ref-iterations", "10", "Refinement iterations"},
    {"warps", "2", "Refinement warps"},
    {"precision", "fp32", "Feature CNN arithmetic on MI355X: fp32 (as the reference) or bf16 (MFMA, faster)"},
    {"cnn-gc", "true", "Free each feature map once consumed (false: keep them and run the two towers on two streams)"},
};

const Spec* find(const std::string& key) {
  for (const Spec& s : kSpecs)
    if (key == s.key) return &s;
  return nullptr;
}

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

[[noreturn]] void opt_fail(const std::string& msg) { throw Error(CCS_ERR_ARG, msg); }

}  // namespace

Options::Options() {
  for (const Spec& s : kSpecs)
    if (s.def) values_[s.key] = s.def;
}

void Options::set(const std::string& key, const std::string& value, bool from_cli) {
  if (!find(key)) opt_fail("unknown option '" + key + "'");
  if (!from_cli && explicit_.count(key)) return;  // the command line wins
  values_[key] = value;
  if (from_cli) explicit_[key] = true;
}

void Options::parseFile(const std::string& path, bool override_existing) {
  std::ifstream f(path);
  if (!f) opt_fail("cannot read config file " + path);
  std::string line;
  int ln = 0;
  while (std::getline(f, line)) {
    ++ln;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    line = trim(line);
    if (line.empty() || line[0] == '[') continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) opt_fail(path + ":" + std::to_string(ln) + ": expected key = value");
    const std::string key = trim(line.substr(0, eq)), value = trim(line.substr(eq + 1));
    if (!find(key)) opt_fail(path + ":" + std::to_string(ln) + ": unknown option '" + key + "'");
    if (override_existing || !explicit_.count(key)) values_[key] = value;
  }
}

void Options::parse(int argc, const char* const* argv) {
  int positional = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
      std::string key = a.substr(2), value;
      const size_t eq = key.find('=');
      if (eq != std::string::npos) {
        value = key.substr(eq + 1);
        key.resize(eq);
      } else if (key == "help") {
        value = "1";
      } else {
        if (i + 1 >= argc) opt_fail("option --" + key + " needs a value");
        value = argv[++i];
      }
      set(key, value, true);
    } else if (a == "-o" || a == "-g") {
      if (i + 1 >= argc) opt_fail("option " + a + " needs a value");
      set(a == "-o" ? "output-file" : "gpu", argv[++i], true);
    } else if (!a.empty() && a[0] == '-' && a.size() > 1 && !std::isdigit((unsigned char)a[1])) {
      opt_fail("unknown option '" + a + "'");
    } else {
      if (positional >= 2) opt_fail("too many positional arguments");
      set(positional == 0 ? "im0" : "im1", a, true);
      ++positional;
    }
  }
  if (has("config-file")) parseFile(str("config-file"), false);
}

bool Options::has(const std::string& key) const { return values_.count(key) > 0; }

std::string Options::str(const std::string& key) const {
  auto it = values_.find(key);
  if (it == values_.end()) opt_fail("option '" + key + "' is not set");
  return it->second;
}

double Options::num(const std::string& key) const {
  const std::string v = str(key);
  char* end = nullptr;
  const double d = std::strtod(v.c_str(), &end);
  if (v.empty() || *end != '\0') opt_fail("option '" + key + "': not a number: '" + v + "'");
  return d;
}

int Options::integer(const std::string& key) const {
  const double d = num(key);
  if (d != (double)(int)d) opt_fail("option '" + key + "': not an integer");
  return (int)d;
}

bool Options::flag(const std::string& key) const {
  const std::string v = str(key);
  if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
  if (v == "0" || v == "false" || v == "no" || v == "off") return false;
  opt_fail("option '" + key + "': not a boolean: '" + v + "'");
}

std::string Options::usage() {
  std::ostringstream o;
  o << "Usage: stereo_img [options] im0 im1\n";
  for (const Spec& s : kSpecs) {
    o << "  --" << s.key;
    if (s.def) o << " (=" << (*s.def ? s.def : "\"\"") << ")";
    o << "\t" << s.help << "\n";
  }
  return o.str();
}

}  // namespace ccs
