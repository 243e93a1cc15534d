// ITK MATLAB Level-4 transform files (0GenericAffine.mat), as crates/fastwarp-reg/src/transform/
// itk_mat.rs reads and writes them (after ITK's itkMatlabTransformIO): a transform is two
// variables, the parameters named after the transform class (`AffineTransform_double_3_3`: the
// matrix row-major, then the translation) and `fixed` (the centre). Each variable is a 20-byte
// header of five little-endian u32 (type: 0 for f64, 10 for f32; rows; cols = 1; imaginary = 0;
// name length with its NUL), the NUL-terminated name, then `rows` values. An affine file is
// 193 bytes.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "common.hpp"
#include "transform.hpp"

namespace fw {

constexpr const char* kItkAffineName = "AffineTransform_double_3_3";

namespace detail {

inline void put_u32(std::vector<uint8_t>& out, uint32_t v) {
  for (int b = 0; b < 4; ++b) out.push_back(uint8_t(v >> (8 * b)));
}

inline void put_variable(std::vector<uint8_t>& out, const std::string& name, const double* values, uint32_t n) {
  put_u32(out, 0);  // f64
  put_u32(out, n);
  put_u32(out, 1);
  put_u32(out, 0);
  put_u32(out, uint32_t(name.size() + 1));
  out.insert(out.end(), name.begin(), name.end());
  out.push_back(0);
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t bits;
    std::memcpy(&bits, &values[i], 8);
    for (int b = 0; b < 8; ++b) out.push_back(uint8_t(bits >> (8 * b)));
  }
}

inline uint32_t get_u32(const std::vector<uint8_t>& b, size_t pos) {
  if (pos + 4 > b.size()) invalid("ITK MATLAB transform (.mat)", "truncated header");
  return uint32_t(b[pos]) | uint32_t(b[pos + 1]) << 8 | uint32_t(b[pos + 2]) << 16 | uint32_t(b[pos + 3]) << 24;
}

struct MatVariable {
  std::string name;
  std::vector<double> values;
};

// itk_mat.rs:52-99 `read_variable`: f64 or f32 column vectors only; the name ends at its first NUL.
inline MatVariable get_variable(const std::vector<uint8_t>& b, size_t& pos) {
  const char* what = "ITK MATLAB transform (.mat)";
  const uint32_t type = get_u32(b, pos), rows = get_u32(b, pos + 4), cols = get_u32(b, pos + 8);
  const uint32_t imag = get_u32(b, pos + 12), namelen = get_u32(b, pos + 16);
  if (imag != 0) invalid(what, "complex variables are not supported");
  if (cols != 1) invalid(what, "only column-vector parameters are supported");
  if (type != 0 && type != 10) invalid(what, "unsupported MATLAB v4 type code " + std::to_string(type));
  const size_t width = type == 0 ? 8 : 4;
  const size_t name_start = pos + 20, data_start = name_start + namelen;
  const size_t data_end = data_start + size_t(rows) * width;
  if (namelen == 0 || data_end > b.size()) invalid(what, "truncated variable");
  MatVariable v;
  for (size_t i = name_start; i < data_start && b[i] != 0; ++i) v.name.push_back(char(b[i]));
  for (size_t i = 0; i < rows; ++i) {
    const size_t at = data_start + i * width;
    if (type == 0) {
      uint64_t bits = 0;
      for (int k = 0; k < 8; ++k) bits |= uint64_t(b[at + size_t(k)]) << (8 * k);
      double d;
      std::memcpy(&d, &bits, 8);
      v.values.push_back(d);
    } else {
      uint32_t bits = 0;
      for (int k = 0; k < 4; ++k) bits |= uint32_t(b[at + size_t(k)]) << (8 * k);
      float f;
      std::memcpy(&f, &bits, 4);
      v.values.push_back(double(f));
    }
  }
  pos = data_end;
  return v;
}

}  // namespace detail

// itk_mat.rs:150-165 `format_mat` for one affine transform.
inline std::vector<uint8_t> format_affine_mat(const AffineTransform& t) {
  double params[12], fixed[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) params[3 * r + c] = t.matrix[r][c];
    params[9 + r] = t.translation[r];
    fixed[r] = t.center[r];
  }
  std::vector<uint8_t> out;
  detail::put_variable(out, kItkAffineName, params, 12);
  detail::put_variable(out, "fixed", fixed, 3);
  return out;
}

// itk_mat.rs:102-133 `parse_mat` with itk_params.rs:58-102 `elem_from_record`, for the first
// transform of a file of affine transforms (an ANTs affine file holds exactly one): the 3D double
// or float forms of AffineTransform, MatrixOffsetTransformBase and Rigid3DTransform (12 parameters)
// and TranslationTransform (3). Other classes are refused.
inline AffineTransform parse_affine_mat(const std::vector<uint8_t>& b) {
  const char* what = "ITK MATLAB transform (.mat)";
  if (b.empty()) invalid(what, "no transform in file");
  size_t pos = 0;
  const detail::MatVariable params = detail::get_variable(b, pos);
  if (pos >= b.size()) invalid(what, "missing `fixed` variable after the parameters");
  const detail::MatVariable fixed = detail::get_variable(b, pos);
  if (fixed.name != "fixed") invalid(what, "expected variable `fixed`, found `" + fixed.name + "`");
  std::string base = params.name;
  for (const char* suffix : {"_double_3_3", "_float_3_3"})
    if (base.size() > std::strlen(suffix) && base.compare(base.size() - std::strlen(suffix), std::string::npos, suffix) == 0) {
      base.resize(base.size() - std::strlen(suffix));
      break;
    }
  AffineTransform t;
  if (base == "TranslationTransform") {
    if (params.values.size() != 3) invalid(what, "translation parameters: expected 3 values");
    for (int r = 0; r < 3; ++r) t.translation[r] = params.values[size_t(r)];
    return t;
  }
  if (base != "AffineTransform" && base != "MatrixOffsetTransformBase" && base != "Rigid3DTransform")
    invalid(what, "unsupported transform type " + params.name);
  if (params.values.size() != 12) invalid(what, "affine parameters: expected 12 values");
  if (fixed.values.size() != 3) invalid(what, "affine fixed parameters: expected 3 values");
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) t.matrix[r][c] = params.values[size_t(3 * r + c)];
    t.translation[r] = params.values[size_t(9 + r)];
    t.center[r] = fixed.values[size_t(r)];
  }
  return t;
}

inline void write_affine_mat(const std::string& path, const AffineTransform& t) {
  const std::vector<uint8_t> bytes = format_affine_mat(t);
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) invalid("io", "cannot open " + path + " for writing");
  f.write(reinterpret_cast<const char*>(bytes.data()), std::streamsize(bytes.size()));
  if (!f) invalid("io", "cannot write " + path);
}

inline AffineTransform read_affine_mat(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) invalid("io", "cannot open " + path);
  const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return parse_affine_mat(bytes);
}

}  // namespace fw
