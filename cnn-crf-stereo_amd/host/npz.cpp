// .npz reader/writer for the parameter files (replaces cnpy, which the
// reference's Operator::setParams uses: stereo/operators/operator.cpp:55-85).
// A .npz is a ZIP archive whose members are .npy files; numpy.savez stores
// them uncompressed, numpy.savez_compressed deflates them (zlib, raw deflate).
#include <zlib.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/ccs_stereo.hpp"

namespace ccs {
namespace {

[[noreturn]] void io_fail(const std::string& msg) { throw Error(CCS_ERR_IO, msg); }

uint16_t rd16(const unsigned char* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const unsigned char* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

std::vector<unsigned char> read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) io_fail("cannot open " + path);
  return std::vector<unsigned char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// Parses one .npy image (float32 or float64, C order) into a float array.
HostArray parse_npy(const unsigned char* p, size_t n, const std::string& name) {
  if (n < 10 || std::memcmp(p, "\x93NUMPY", 6) != 0) io_fail(name + ": not a .npy member");
  const int major = p[6];
  size_t hlen, off;
  if (major == 1) {
    hlen = rd16(p + 8);
    off = 10;
  } else if (major == 2 || major == 3) {
    if (n < 12) io_fail(name + ": truncated header");
    hlen = rd32(p + 8);
    off = 12;
  } else {
    io_fail(name + ": unsupported .npy version");
  }
  if (off + hlen > n) io_fail(name + ": truncated header");
  const std::string hdr((const char*)p + off, hlen);
  auto field = [&](const char* key) -> std::string {
    const size_t k = hdr.find(key);
    if (k == std::string::npos) io_fail(name + ": header lacks " + key);
    return hdr.substr(k + std::strlen(key));
  };
  std::string descr = field("'descr':");
  const size_t q0 = descr.find('\''), q1 = descr.find('\'', q0 + 1);
  descr = descr.substr(q0 + 1, q1 - q0 - 1);
  if (field("'fortran_order':").find("False") > 2) io_fail(name + ": Fortran-order arrays are not supported");
  std::string shp = field("'shape':");
  shp = shp.substr(shp.find('(') + 1, shp.find(')') - shp.find('(') - 1);
  HostArray a;
  std::stringstream ss(shp);
  std::string tok;
  while (std::getline(ss, tok, ',')) {
    if (tok.find_first_not_of(" ") == std::string::npos) continue;
    a.shape.push_back(std::stoi(tok));
  }
  const size_t count = a.size();
  const unsigned char* data = p + off + hlen;
  const size_t avail = n - off - hlen;
  a.data.resize(count);
  if (descr == "<f4") {
    if (avail < count * 4) io_fail(name + ": truncated data");
    std::memcpy(a.data.data(), data, count * 4);
  } else if (descr == "<f8") {
    if (avail < count * 8) io_fail(name + ": truncated data");
    for (size_t i = 0; i < count; ++i) {
      double v;
      std::memcpy(&v, data + 8 * i, 8);
      a.data[i] = (float)v;
    }
  } else {
    io_fail(name + ": dtype " + descr + " (need <f4 or <f8)");
  }
  return a;
}

std::vector<unsigned char> inflate_raw(const unsigned char* src, size_t n, size_t out_size, const std::string& name) {
  std::vector<unsigned char> out(out_size);
  z_stream zs{};
  if (inflateInit2(&zs, -MAX_WBITS) != Z_OK) io_fail(name + ": inflateInit failed");
  zs.next_in = const_cast<unsigned char*>(src);
  zs.avail_in = (uInt)n;
  zs.next_out = out.data();
  zs.avail_out = (uInt)out_size;
  const int rc = inflate(&zs, Z_FINISH);
  inflateEnd(&zs);
  if (rc != Z_STREAM_END || zs.total_out != out_size) io_fail(name + ": corrupt deflate stream");
  return out;
}

// npy v1.0 image of a float32 array: magic, version, header length, the header
// dict padded with spaces to 64-byte alignment and ended by a newline, data.
std::vector<unsigned char> make_npy(const HostArray& a, const std::string& key) {
  if (a.data.size() != a.size()) throw Error(CCS_ERR_ARG, "npz: data size does not match the shape of " + key);
  std::string dict = "{'descr': '<f4', 'fortran_order': False, 'shape': (";
  for (size_t i = 0; i < a.shape.size(); ++i)
    dict += std::to_string(a.shape[i]) + (a.shape.size() == 1 ? "," : (i + 1 < a.shape.size() ? ", " : ""));
  dict += "), }";
  const size_t total = 10 + dict.size() + 1;
  dict.append((64 - total % 64) % 64, ' ');
  dict += '\n';
  std::vector<unsigned char> npy = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  npy.push_back(dict.size() & 0xff);
  npy.push_back((dict.size() >> 8) & 0xff);
  npy.insert(npy.end(), dict.begin(), dict.end());
  const unsigned char* d = reinterpret_cast<const unsigned char*>(a.data.data());
  npy.insert(npy.end(), d, d + a.data.size() * 4);
  return npy;
}

}  // namespace

namespace npz {

void save_npy(const std::string& path, const HostArray& a) {
  const std::vector<unsigned char> npy = make_npy(a, path);
  std::ofstream f(path, std::ios::binary);
  if (!f) io_fail("cannot write " + path);
  f.write((const char*)npy.data(), (std::streamsize)npy.size());
  if (!f) io_fail("write failed: " + path);
}

std::map<std::string, HostArray> load(const std::string& path) {
  const std::vector<unsigned char> z = read_file(path);
  const size_t n = z.size();
  if (n < 22) io_fail(path + ": not a ZIP archive");
  // End of central directory: last occurrence of its signature.
  size_t eocd = std::string::npos;
  for (size_t i = n - 22 + 1; i-- > 0 && n - i <= 22 + 65535;)
    if (rd32(&z[i]) == 0x06054b50u) {
      eocd = i;
      break;
    }
  if (eocd == std::string::npos) io_fail(path + ": no ZIP end record");
  const unsigned entries = rd16(&z[eocd + 10]);
  size_t cd = rd32(&z[eocd + 16]);
  std::map<std::string, HostArray> out;
  for (unsigned e = 0; e < entries; ++e) {
    if (cd + 46 > n || rd32(&z[cd]) != 0x02014b50u) io_fail(path + ": corrupt central directory");
    const unsigned method = rd16(&z[cd + 10]);
    const uint32_t csize = rd32(&z[cd + 20]), usize = rd32(&z[cd + 24]);
    const unsigned nlen = rd16(&z[cd + 28]), xlen = rd16(&z[cd + 30]), clen = rd16(&z[cd + 32]);
    const uint32_t loff = rd32(&z[cd + 42]);
    std::string name((const char*)&z[cd + 46], nlen);
    if (csize == 0xffffffffu || usize == 0xffffffffu || loff == 0xffffffffu)
      io_fail(path + ": ZIP64 members are not supported (" + name + ")");
    if (loff + 30 > n || rd32(&z[loff]) != 0x04034b50u) io_fail(path + ": corrupt local header");
    const size_t data = loff + 30 + rd16(&z[loff + 26]) + rd16(&z[loff + 28]);
    if (data + csize > n) io_fail(path + ": truncated member " + name);
    if (name.size() > 4 && name.compare(name.size() - 4, 4, ".npy") == 0) name.resize(name.size() - 4);
    if (method == 0) {
      out[name] = parse_npy(&z[data], csize, name);
    } else if (method == 8) {
      const std::vector<unsigned char> raw = inflate_raw(&z[data], csize, usize, name);
      out[name] = parse_npy(raw.data(), raw.size(), name);
    } else {
      io_fail(path + ": member " + name + " uses unsupported compression");
    }
    cd += 46 + nlen + xlen + clen;
  }
  return out;
}

HostArray load_array(const std::string& path, const std::string& key) {
  std::map<std::string, HostArray> all = load(path);
  auto it = all.find(key);
  if (it == all.end()) io_fail(path + ": no array '" + key + "'");
  return std::move(it->second);
}

HostArray load_npy(const std::string& path) {
  const std::vector<unsigned char> b = read_file(path);
  return parse_npy(b.data(), b.size(), path);
}

void save(const std::string& path, const std::vector<std::pair<std::string, HostArray>>& arrays) {
  std::vector<unsigned char> zip, cdir;
  auto put16 = [](std::vector<unsigned char>& v, uint32_t x) {
    v.push_back(x & 0xff);
    v.push_back((x >> 8) & 0xff);
  };
  auto put32 = [&](std::vector<unsigned char>& v, uint32_t x) {
    put16(v, x & 0xffff);
    put16(v, x >> 16);
  };
  for (const auto& [key, a] : arrays) {
    const std::vector<unsigned char> npy = make_npy(a, key);
    const std::string name = key + ".npy";
    const uint32_t crc = (uint32_t)crc32(0L, npy.data(), (uInt)npy.size());
    const uint32_t off = (uint32_t)zip.size();
    put32(zip, 0x04034b50u);
    put16(zip, 20); put16(zip, 0); put16(zip, 0); put16(zip, 0); put16(zip, 0x21);
    put32(zip, crc); put32(zip, (uint32_t)npy.size()); put32(zip, (uint32_t)npy.size());
    put16(zip, (uint32_t)name.size()); put16(zip, 0);
    zip.insert(zip.end(), name.begin(), name.end());
    zip.insert(zip.end(), npy.begin(), npy.end());
    put32(cdir, 0x02014b50u);
    put16(cdir, 20); put16(cdir, 20); put16(cdir, 0); put16(cdir, 0); put16(cdir, 0); put16(cdir, 0x21);
    put32(cdir, crc); put32(cdir, (uint32_t)npy.size()); put32(cdir, (uint32_t)npy.size());
    put16(cdir, (uint32_t)name.size()); put16(cdir, 0); put16(cdir, 0); put16(cdir, 0); put16(cdir, 0);
    put32(cdir, 0); put32(cdir, off);
    cdir.insert(cdir.end(), name.begin(), name.end());
  }
  const uint32_t cd_off = (uint32_t)zip.size();
  zip.insert(zip.end(), cdir.begin(), cdir.end());
  put32(zip, 0x06054b50u);
  put16(zip, 0); put16(zip, 0);
  put16(zip, (uint32_t)arrays.size()); put16(zip, (uint32_t)arrays.size());
  put32(zip, (uint32_t)cdir.size()); put32(zip, cd_off); put16(zip, 0);
  std::ofstream f(path, std::ios::binary);
  if (!f) io_fail("cannot write " + path);
  f.write((const char*)zip.data(), (std::streamsize)zip.size());
  if (!f) io_fail("write failed: " + path);
}

}  // namespace npz
}  // namespace ccs
