// Image files for stereo_img: PNG (zlib), binary PGM/PPM and PFM readers, PFM
// and 16-bit PNG writers (the formats of stereo/tools.cpp:29-62: PFM rows are
// stored bottom-up with scale -1 (little endian); the PNG holds disparity * 256
// as 16-bit gray, clamped at 0).  Images load as RGBA float in [0, 1].
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/ccs_stereo.hpp"

namespace ccs {
namespace {

[[noreturn]] void io_fail(const std::string& msg) { throw Error(CCS_ERR_IO, msg); }

std::vector<unsigned char> read_all(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) io_fail("cannot open " + path);
  return std::vector<unsigned char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

uint32_t be32(const unsigned char* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

Image decode_png(const std::vector<unsigned char>& f, const std::string& path) {
  if (f.size() < 8 || std::memcmp(f.data(), "\x89PNG\r\n\x1a\n", 8) != 0) io_fail(path + ": not a PNG file");
  size_t pos = 8;
  uint32_t W = 0, H = 0;
  int depth = 0, ctype = -1;
  std::vector<unsigned char> idat, palette;
  while (pos + 12 <= f.size()) {
    const uint32_t len = be32(&f[pos]);
    const std::string type((const char*)&f[pos + 4], 4);
    if (pos + 12 + (size_t)len > f.size()) io_fail(path + ": truncated chunk");
    const unsigned char* d = &f[pos + 8];
    if (type == "IHDR") {
      W = be32(d);
      H = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      if (d[12] != 0) io_fail(path + ": interlaced PNG is not supported");
    } else if (type == "PLTE") {
      palette.assign(d, d + len);
    } else if (type == "IDAT") {
      idat.insert(idat.end(), d, d + len);
    } else if (type == "IEND") {
      break;
    }
    pos += 12 + len;
  }
  const int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
  if (!W || !H || !ch || (depth != 8 && depth != 16) || (ctype == 3 && depth != 8))
    io_fail(path + ": unsupported PNG (need 8/16-bit gray, RGB, gray+alpha, RGBA or 8-bit palette)");
  const size_t bpp = (size_t)ch * depth / 8, stride = bpp * W;
  std::vector<unsigned char> raw((stride + 1) * H);
  uLongf out_len = (uLongf)raw.size();
  if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
    io_fail(path + ": corrupt image data");
  std::vector<unsigned char> px(stride * H), prev(stride, 0);
  for (uint32_t y = 0; y < H; ++y) {
    const unsigned char* r = &raw[y * (stride + 1)];
    unsigned char* o = &px[y * stride];
    const int ft = r[0];
    for (size_t x = 0; x < stride; ++x) {
      const int a = x >= bpp ? o[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
      int v = r[1 + x];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) / 2; break;
        case 4: v += paeth(a, b, c); break;
        default: io_fail(path + ": bad filter type");
      }
      o[x] = (unsigned char)v;
    }
    std::copy(o, o + stride, prev.begin());
  }
  Image img;
  img.width = (int)W;
  img.height = (int)H;
  img.rgba.assign((size_t)W * H * 4, 1.0f);
  const float scale = depth == 16 ? 1.0f / 65535.f : 1.0f / 255.f;
  for (size_t i = 0; i < (size_t)W * H; ++i) {
    float s[4];
    for (int c = 0; c < ch; ++c)
      s[c] = depth == 16 ? (float)((px[i * bpp + 2 * c] << 8) | px[i * bpp + 2 * c + 1]) : (float)px[i * bpp + c];
    float* o = &img.rgba[4 * i];
    if (ctype == 3) {
      const size_t k = (size_t)s[0] * 3;
      if (k + 2 >= palette.size()) io_fail(path + ": palette index out of range");
      o[0] = palette[k] / 255.f;
      o[1] = palette[k + 1] / 255.f;
      o[2] = palette[k + 2] / 255.f;
    } else if (ch <= 2) {
      o[0] = o[1] = o[2] = s[0] * scale;
      if (ch == 2) o[3] = s[1] * scale;
    } else {
      for (int c = 0; c < ch; ++c) o[c] = s[c] * scale;
    }
  }
  return img;
}

// Binary PGM (P5) / PPM (P6), 8 or 16 bit.
Image decode_pnm(const std::vector<unsigned char>& f, const std::string& path) {
  std::string hdr;
  size_t pos = 0;
  int fields = 0, vals[3] = {0, 0, 0};
  const bool color = f[1] == '6';
  pos = 2;
  while (fields < 3 && pos < f.size()) {
    while (pos < f.size() && std::isspace(f[pos])) ++pos;
    if (pos < f.size() && f[pos] == '#') {
      while (pos < f.size() && f[pos] != '\n') ++pos;
      continue;
    }
    int v = 0;
    bool any = false;
    while (pos < f.size() && std::isdigit(f[pos])) {
      v = v * 10 + (f[pos++] - '0');
      any = true;
    }
    if (!any) io_fail(path + ": bad PNM header");
    vals[fields++] = v;
  }
  ++pos;  // single whitespace after maxval
  const int W = vals[0], H = vals[1], maxv = vals[2];
  const int ch = color ? 3 : 1, bps = maxv > 255 ? 2 : 1;
  if (W <= 0 || H <= 0 || maxv <= 0 || pos + (size_t)W * H * ch * bps > f.size()) io_fail(path + ": truncated PNM");
  Image img;
  img.width = W;
  img.height = H;
  img.rgba.assign((size_t)W * H * 4, 1.0f);
  for (size_t i = 0; i < (size_t)W * H; ++i)
    for (int c = 0; c < ch; ++c) {
      const size_t k = pos + (i * ch + c) * bps;
      const float v = (bps == 2 ? (float)((f[k] << 8) | f[k + 1]) : (float)f[k]) / (float)maxv;
      if (color) img.rgba[4 * i + c] = v;
      else img.rgba[4 * i] = img.rgba[4 * i + 1] = img.rgba[4 * i + 2] = v;
    }
  return img;
}

void put_be32(std::vector<unsigned char>& v, uint32_t x) {
  for (int s = 24; s >= 0; s -= 8) v.push_back((x >> s) & 0xff);
}

void png_chunk(std::ofstream& o, const char* type, const std::vector<unsigned char>& data) {
  std::vector<unsigned char> buf;
  put_be32(buf, (uint32_t)data.size());
  buf.insert(buf.end(), type, type + 4);
  buf.insert(buf.end(), data.begin(), data.end());
  const uint32_t crc = (uint32_t)crc32(0L, buf.data() + 4, (uInt)(buf.size() - 4));
  put_be32(buf, crc);
  o.write((const char*)buf.data(), (std::streamsize)buf.size());
}

}  // namespace

Image imread_c4(const std::string& path) {
  const std::vector<unsigned char> f = read_all(path);
  if (f.size() >= 8 && f[0] == 0x89 && f[1] == 'P') return decode_png(f, path);
  if (f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return decode_pnm(f, path);
  if (f.size() >= 2 && f[0] == 'P' && (f[1] == 'f' || f[1] == 'F')) {
    int w = 0, h = 0;
    std::vector<float> v = load_pfm(path, &w, &h);
    Image img;
    img.width = w;
    img.height = h;
    img.rgba.assign((size_t)w * h * 4, 1.0f);
    const int ch = f[1] == 'F' ? 3 : 1;
    for (size_t i = 0; i < (size_t)w * h; ++i)
      for (int c = 0; c < 3; ++c) img.rgba[4 * i + c] = v[i * ch + (ch == 3 ? c : 0)];
    return img;
  }
  io_fail(path + ": unknown image format (PNG, PGM/PPM or PFM)");
}

std::vector<float> load_pfm(const std::string& path, int* width, int* height) {
  const std::vector<unsigned char> f = read_all(path);
  std::string text((const char*)f.data(), std::min<size_t>(f.size(), 256));
  std::istringstream is(text);
  std::string magic;
  double scale = 0;
  int w = 0, h = 0;
  is >> magic >> w >> h >> scale;
  if ((magic != "Pf" && magic != "PF") || w <= 0 || h <= 0 || scale == 0) io_fail(path + ": bad PFM header");
  const size_t off = (size_t)is.tellg() + 1;
  const int ch = magic == "PF" ? 3 : 1;
  const size_t n = (size_t)w * h * ch;
  if (off + 4 * n > f.size()) io_fail(path + ": truncated PFM");
  if (scale > 0) io_fail(path + ": big-endian PFM is not supported");
  std::vector<float> out(n);
  for (int y = 0; y < h; ++y)  // stored bottom-up
    std::memcpy(&out[(size_t)(h - 1 - y) * w * ch], &f[off + (size_t)y * w * ch * 4], (size_t)w * ch * 4);
  *width = w;
  *height = h;
  return out;
}

void save_pfm(const std::string& path, const float* data, int width, int height) {
  std::ofstream o(path, std::ios::binary);
  if (!o) io_fail("cannot write " + path);
  o << "Pf\n" << width << " " << height << "\n-1.0\n";
  for (int y = height - 1; y >= 0; --y) o.write((const char*)(data + (size_t)y * width), (std::streamsize)width * 4);
  if (!o) io_fail("write failed: " + path);
}

void save_png_disparity(const std::string& path, const float* data, int width, int height, bool kitti) {
  float mx = 0.f;
  if (!kitti)
    for (size_t i = 0; i < (size_t)width * height; ++i) mx = std::max(mx, data[i]);
  const int depth = kitti ? 16 : 8;
  const size_t stride = (size_t)width * depth / 8;
  std::vector<unsigned char> raw((stride + 1) * height);
  for (int y = 0; y < height; ++y) {
    unsigned char* r = &raw[y * (stride + 1)];
    r[0] = 0;
    for (int x = 0; x < width; ++x) {
      const float v = data[(size_t)y * width + x];
      if (kitti) {
        const double s = std::min(65535.0, std::max(v * 256.0, 0.0));
        const uint16_t q = (uint16_t)s;
        r[1 + 2 * x] = q >> 8;
        r[2 + 2 * x] = q & 0xff;
      } else {
        r[1 + x] = (unsigned char)(mx > 0 ? std::lround(std::max(0.f, v) / mx * 255.f) : 0);
      }
    }
  }
  uLongf clen = compressBound((uLong)raw.size());
  std::vector<unsigned char> comp(clen);
  if (compress2(comp.data(), &clen, raw.data(), (uLong)raw.size(), 6) != Z_OK) io_fail("PNG compression failed");
  comp.resize(clen);
  std::ofstream o(path, std::ios::binary);
  if (!o) io_fail("cannot write " + path);
  o.write("\x89PNG\r\n\x1a\n", 8);
  std::vector<unsigned char> ihdr;
  put_be32(ihdr, (uint32_t)width);
  put_be32(ihdr, (uint32_t)height);
  ihdr.insert(ihdr.end(), {(unsigned char)depth, 0, 0, 0, 0});
  png_chunk(o, "IHDR", ihdr);
  png_chunk(o, "IDAT", comp);
  png_chunk(o, "IEND", {});
  if (!o) io_fail("write failed: " + path);
}

}  // namespace ccs
