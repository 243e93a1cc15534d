// ITK HDF5 transform files: the Composite.h5 / InverseComposite.h5 that antsRegistration writes
// (crates/fastwarp-reg/src/transform/itk_h5.rs, ants_io.rs; the layout of ITK's
// itkHDF5TransformIO). The root holds the strings HDFVersion, ITKVersion, OSName and OSVersion and
// the group TransformGroup, whose member "0" holds only TransformType =
// CompositeTransform_double_3_3 and whose members "1", "2", ... hold the chain in queue order,
// each with TransformType, TransformFixedParameters and TransformParameters (f64 datasets).
// An affine is AffineTransform_double_3_3: 12 parameters (matrix row-major, then translation)
// and its centre as the 3 fixed parameters. A displacement field is
// DisplacementFieldTransform_double_3_3: 18 fixed parameters (size, origin, spacing, direction
// row-major) and 3 N parameters, the displacement of each voxel (x, y, z) with x fastest.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "grid.hpp"
#include "hdf5.hpp"
#include "transform.hpp"

namespace fw {

// One element of a transform chain: an affine, or a displacement field held in f32 (as the
// reference's DisplacementField holds it), 3 values per voxel.
struct ChainElement {
  bool is_field = false;
  AffineTransform affine;
  Grid grid;
  std::vector<float> field;
};

namespace itk_h5 {

constexpr const char* kWhat = "ITK HDF5 transform";
constexpr const char* kGroup = "TransformGroup";
constexpr const char* kComposite = "CompositeTransform_double_3_3";
constexpr const char* kAffine = "AffineTransform_double_3_3";
constexpr const char* kField = "DisplacementFieldTransform_double_3_3";
// itk_h5.rs:47-51: the library and toolkit versions the layout reproduces
constexpr const char* kHdfVersion = "1.14.0";
constexpr const char* kItkVersion = "5.4.4";

// grid.rs:158-169 `fixed_parameters`
inline std::vector<double> fixed_parameters(const Grid& g) {
  std::vector<double> p(18);
  for (int a = 0; a < 3; ++a) {
    p[size_t(a)] = double(g.size[size_t(a)]);
    p[size_t(3 + a)] = g.origin[a];
    p[size_t(6 + a)] = g.spacing[a];
    for (int b = 0; b < 3; ++b) p[size_t(9 + 3 * a + b)] = g.direction[a][b];
  }
  return p;
}

// grid.rs:172-195 `from_fixed_parameters`: whole positive sizes, then the grid's own checks.
inline Grid grid_from_fixed(const std::vector<double>& p) {
  if (p.size() != 18) invalid(kWhat, "displacement field fixed parameters: expected 18 values, got " + std::to_string(p.size()));
  std::array<size_t, 3> size{};
  for (int a = 0; a < 3; ++a) {
    const double v = p[size_t(a)];
    if (!(std::isfinite(v) && v >= 1.0 && v == std::floor(v) && v <= double(size_t(1) << 20)))
      invalid(kWhat, "bad size parameter " + std::to_string(v));
    size[size_t(a)] = size_t(v);
  }
  Mat3 dir{};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) dir[a][b] = p[size_t(9 + 3 * a + b)];
  return Grid(size, {p[6], p[7], p[8]}, {p[3], p[4], p[5]}, dir);
}

// The class of a 3D type name without its _double_3_3 / _float_3_3 suffix ("" for other names).
inline std::string base_type(const std::string& name) {
  for (const std::string suffix : {"_double_3_3", "_float_3_3"})
    if (name.size() > suffix.size() && name.compare(name.size() - suffix.size(), suffix.size(), suffix) == 0)
      return name.substr(0, name.size() - suffix.size());
  return "";
}

// itk_params.rs:58-102 `elem_from_record`: the composite itself gives no element.
inline bool element_from_record(const std::string& type, const std::vector<double>& params,
                                const std::vector<double>& fixed, ChainElement& out) {
  const std::string base = base_type(type);
  if (base.empty()) invalid(kWhat, "transform type " + type + " (only 3D double or float types)");
  if (base == "CompositeTransform") return false;
  out = ChainElement{};
  if (base == "AffineTransform" || base == "MatrixOffsetTransformBase" || base == "Rigid3DTransform") {
    if (params.size() != 12) invalid(kWhat, "affine parameters: expected 12 values");
    if (fixed.size() != 3) invalid(kWhat, "affine fixed parameters: expected 3 values");
    out.affine = AffineTransform::from_parameters12(params.data(), {fixed[0], fixed[1], fixed[2]});
  } else if (base == "Euler3DTransform") {
    if (params.size() != 6) invalid(kWhat, "Euler3D parameters: expected 6 values");
    if (fixed.size() != 3 && fixed.size() != 4) invalid(kWhat, "Euler3D fixed parameters must have 3 or 4 values");
    out.affine = AffineTransform::from_euler3d(params.data(), {fixed[0], fixed[1], fixed[2]}, fixed.size() == 4 && fixed[3] != 0.0);
  } else if (base == "TranslationTransform") {
    if (params.size() != 3) invalid(kWhat, "translation parameters: expected 3 values");
    out.affine = AffineTransform::from_translation({params[0], params[1], params[2]});
  } else if (base == "DisplacementFieldTransform") {
    out.is_field = true;
    out.grid = grid_from_fixed(fixed);
    if (params.size() != 3 * out.grid.num_voxels())
      invalid(kWhat, "displacement field parameters: expected 3 values per voxel");
    out.field.resize(params.size());
    for (size_t i = 0; i < params.size(); ++i) out.field[i] = float(params[i]);
  } else {
    invalid(kWhat, "transform class " + base);
  }
  return true;
}

// The OS strings ITK records; the reference writes its OS and architecture names.
inline std::string os_name() {
#if defined(__linux__)
  return "linux";
#else
  return "unknown";
#endif
}
inline std::string os_version() {
#if defined(__x86_64__)
  return "x86_64";
#elif defined(__aarch64__)
  return "aarch64";
#else
  return "unknown";
#endif
}

}  // namespace itk_h5

// itk_h5.rs:114-199 `format_h5` / `write_h5_file`: group 0 is the composite, groups 1.. the
// chain in queue order. An empty chain is refused.
inline void write_composite_h5(const std::string& path, const std::vector<ChainElement>& chain) {
  using h5::Node;
  if (chain.empty()) invalid(itk_h5::kWhat, "cannot write an empty chain");
  std::vector<Node> groups;
  groups.push_back(Node::group("0", {Node::string("TransformType", itk_h5::kComposite)}));
  for (size_t i = 0; i < chain.size(); ++i) {
    const ChainElement& e = chain[i];
    std::vector<Node> members;
    if (e.is_field) {
      if (e.field.size() != 3 * e.grid.num_voxels()) invalid(itk_h5::kWhat, "a field needs 3 values per voxel of its grid");
      members.push_back(Node::string("TransformType", itk_h5::kField));
      members.push_back(Node::values("TransformFixedParameters", itk_h5::fixed_parameters(e.grid)));
      members.push_back(Node::widened("TransformParameters", e.field.data(), e.field.size()));
    } else {
      const std::array<double, 12> p = e.affine.parameters();
      members.push_back(Node::string("TransformType", itk_h5::kAffine));
      members.push_back(Node::values("TransformFixedParameters", {e.affine.center[0], e.affine.center[1], e.affine.center[2]}));
      members.push_back(Node::values("TransformParameters", std::vector<double>(p.begin(), p.end())));
    }
    groups.push_back(Node::group(std::to_string(i + 1), std::move(members)));
  }
  h5::write_file(path, {Node::string("HDFVersion", itk_h5::kHdfVersion), Node::string("ITKVersion", itk_h5::kItkVersion),
                        Node::string("OSName", itk_h5::os_name()), Node::string("OSVersion", itk_h5::os_version()),
                        Node::group(itk_h5::kGroup, std::move(groups))});
}

// itk_h5.rs:54-110 `read_chain` / `read_h5_file`: the members of TransformGroup in numeric
// order (a non-numeric name is an error); a displacement field keeps its values in f32; a file
// without a transform is an error.
inline std::vector<ChainElement> read_composite_h5(const std::string& path) {
  const h5::File f = h5::File::open(path);
  std::vector<std::pair<unsigned long long, std::string>> names;
  for (const std::string& n : f.members(itk_h5::kGroup)) {
    if (n.empty() || n.find_first_not_of("0123456789") != std::string::npos || n.size() > 19)
      invalid(itk_h5::kWhat, "non-numeric member `" + n + "` in TransformGroup");
    names.emplace_back(std::strtoull(n.c_str(), nullptr, 10), n);
  }
  std::sort(names.begin(), names.end());
  std::vector<ChainElement> chain;
  for (const auto& entry : names) {
    const std::string base = std::string(itk_h5::kGroup) + "/" + entry.second;
    const std::string type = f.read_string(base + "/TransformType");
    std::vector<double> params, fixed;
    if (f.exists(base + "/TransformParameters")) {
      params = f.read_doubles(base + "/TransformParameters");
      fixed = f.read_doubles(base + "/TransformFixedParameters");
    }
    ChainElement e;
    if (itk_h5::element_from_record(type, params, fixed, e)) chain.push_back(std::move(e));
  }
  if (chain.empty()) invalid(itk_h5::kWhat, "HDF5 transform file without a transform");
  return chain;
}

// ants_io.rs:100-148: <prefix>Composite.h5 holds [affine, field] and <prefix>InverseComposite.h5
// [inverse field, affine^-1]; without fields (a linear-only registration), [affine] and
// [affine^-1]. A singular affine is refused before anything is written.
inline void write_composite_pair(const std::string& prefix, const AffineTransform& affine, const Grid* grid,
                                 std::vector<float> field, std::vector<float> inverse_field) {
  const AffineTransform inverse = affine.inverse();
  std::vector<ChainElement> forward(1), backward;
  forward[0].affine = affine;
  if (grid) {
    ChainElement f, b;
    f.is_field = b.is_field = true;
    f.grid = b.grid = *grid;
    f.field = std::move(field);
    b.field = std::move(inverse_field);
    forward.push_back(std::move(f));
    backward.push_back(std::move(b));
  }
  ChainElement a;
  a.affine = inverse;
  backward.push_back(std::move(a));
  write_composite_h5(prefix + "Composite.h5", forward);
  write_composite_h5(prefix + "InverseComposite.h5", backward);
}

}  // namespace fw
