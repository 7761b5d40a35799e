// Minimal .npy reader and writer for the on-disk layout of the relaxation
// loop (simulations/<name>/{input,output}_data/*.npy, io.cu:10-291).
//
// Reading validates the header and converts any of float32, float64, int32,
// int64, uint8 and bool to the requested type, instead of reinterpreting
// the bytes.  Writing produces version 1.0 C-order little-endian arrays.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mepgl {

struct NpyArray {
  std::vector<size_t> shape;
  std::string descr;              // e.g. "<f8"
  bool fortran_order = false;
  std::vector<unsigned char> raw;  // the data bytes as stored

  size_t size() const {
    size_t n = 1;
    for (size_t d : shape) n *= d;
    return n;
  }
  // Values converted to double / int32, in C order.
  std::vector<double> as_double() const;
  std::vector<int32_t> as_int32() const;
};

NpyArray npy_load(const std::string& path);
void npy_save(const std::string& path, const double* data, const std::vector<size_t>& shape);
void npy_save(const std::string& path, const float* data, const std::vector<size_t>& shape);
void npy_save(const std::string& path, const int32_t* data, const std::vector<size_t>& shape);

}  // namespace mepgl
