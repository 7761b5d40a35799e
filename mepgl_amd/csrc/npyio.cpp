// .npy reader and writer (format version 1.0 / 2.0 headers, see npyio.hpp).
#include "npyio.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mepgl {

namespace {

std::string field(const std::string& header, const std::string& key) {
  const size_t k = header.find("'" + key + "'");
  if (k == std::string::npos) throw std::runtime_error("npy header lacks '" + key + "'");
  const size_t colon = header.find(':', k);
  if (colon == std::string::npos) throw std::runtime_error("malformed npy header");
  return header.substr(colon + 1);
}

size_t itemsize(const std::string& d) {
  if (d.size() < 3) throw std::runtime_error("bad dtype " + d);
  return (size_t)std::stoul(d.substr(2));
}

template <class T>
T read_le(const unsigned char* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

template <class OUT>
std::vector<OUT> convert(const NpyArray& a) {
  const size_t n = a.size();
  std::vector<OUT> out(n);
  const std::string& d = a.descr;
  const char kind = d[1];
  const size_t isz = itemsize(d);
  if (d[0] == '>') throw std::runtime_error("big-endian npy arrays are not supported");
  if (a.raw.size() < n * isz) throw std::runtime_error("truncated npy data");
  std::vector<OUT> tmp(n);
  for (size_t i = 0; i < n; ++i) {
    const unsigned char* p = a.raw.data() + i * isz;
    double v;
    if (kind == 'f' && isz == 8) v = read_le<double>(p);
    else if (kind == 'f' && isz == 4) v = read_le<float>(p);
    else if (kind == 'i' && isz == 4) v = read_le<int32_t>(p);
    else if (kind == 'i' && isz == 8) v = (double)read_le<int64_t>(p);
    else if ((kind == 'u' || kind == 'b') && isz == 1) v = *p;
    else throw std::runtime_error("unsupported npy dtype " + d);
    tmp[i] = (OUT)v;
  }
  if (!a.fortran_order || a.shape.size() < 2) return tmp;
  // Fortran order: reorder to C order.
  const size_t nd = a.shape.size();
  std::vector<size_t> idx(nd, 0);
  for (size_t c = 0; c < n; ++c) {
    size_t f = 0, stride = 1;
    for (size_t k = 0; k < nd; ++k) {
      f += idx[k] * stride;
      stride *= a.shape[k];
    }
    out[c] = tmp[f];
    for (size_t k = nd; k-- > 0;) {
      if (++idx[k] < a.shape[k]) break;
      idx[k] = 0;
    }
  }
  return out;
}

void write(const std::string& path, const char* descr, const void* data, size_t isz,
           const std::vector<size_t>& shape) {
  std::ostringstream h;
  h << "{'descr': '" << descr << "', 'fortran_order': False, 'shape': (";
  size_t n = 1;
  for (size_t k = 0; k < shape.size(); ++k) {
    h << shape[k] << (shape.size() == 1 ? "," : (k + 1 < shape.size() ? ", " : ""));
    n *= shape[k];
  }
  h << "), }";
  std::string header = h.str();
  const size_t base = 10;  // magic(6) + version(2) + header length(2)
  size_t total = base + header.size() + 1;
  const size_t pad = (64 - total % 64) % 64;
  header.append(pad, ' ');
  header.push_back('\n');
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  f.write((const char*)magic, 8);
  const uint16_t hl = (uint16_t)header.size();
  f.write((const char*)&hl, 2);
  f.write(header.data(), (std::streamsize)header.size());
  f.write((const char*)data, (std::streamsize)(n * isz));
  if (!f) throw std::runtime_error("write failed: " + path);
}

}  // namespace

NpyArray npy_load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  unsigned char magic[8];
  f.read((char*)magic, 8);
  if (!f || magic[0] != 0x93 || std::memcmp(magic + 1, "NUMPY", 5) != 0)
    throw std::runtime_error(path + " is not an npy file");
  size_t hlen;
  if (magic[6] == 1) {
    uint16_t v;
    f.read((char*)&v, 2);
    hlen = v;
  } else {
    uint32_t v;
    f.read((char*)&v, 4);
    hlen = v;
  }
  std::string header(hlen, '\0');
  f.read(&header[0], (std::streamsize)hlen);
  if (!f) throw std::runtime_error("truncated npy header: " + path);
  NpyArray a;
  {
    std::string d = field(header, "descr");
    const size_t q0 = d.find('\''), q1 = d.find('\'', q0 + 1);
    a.descr = d.substr(q0 + 1, q1 - q0 - 1);
    if (a.descr.size() >= 2 && a.descr[0] == '|') a.descr[0] = '<';
    if (a.descr == "<b1") a.descr = "<b1";
  }
  a.fortran_order = field(header, "fortran_order").find("True") < 8;
  {
    std::string sh = field(header, "shape");
    const size_t p0 = sh.find('('), p1 = sh.find(')');
    std::string body = sh.substr(p0 + 1, p1 - p0 - 1);
    std::stringstream ss(body);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
      size_t b = tok.find_first_not_of(" \t");
      if (b == std::string::npos) continue;
      a.shape.push_back((size_t)std::stoull(tok.substr(b)));
    }
  }
  const size_t bytes = a.size() * itemsize(a.descr);
  a.raw.resize(bytes);
  f.read((char*)a.raw.data(), (std::streamsize)bytes);
  if ((size_t)f.gcount() != bytes) throw std::runtime_error("truncated npy data: " + path);
  return a;
}

std::vector<double> NpyArray::as_double() const { return convert<double>(*this); }
std::vector<int32_t> NpyArray::as_int32() const { return convert<int32_t>(*this); }

void npy_save(const std::string& path, const double* data, const std::vector<size_t>& shape) {
  write(path, "<f8", data, 8, shape);
}
void npy_save(const std::string& path, const float* data, const std::vector<size_t>& shape) {
  write(path, "<f4", data, 4, shape);
}
void npy_save(const std::string& path, const int32_t* data, const std::vector<size_t>& shape) {
  write(path, "<i4", data, 4, shape);
}

}  // namespace mepgl
