// Image grids in ITK conventions, scalar volumes and displacement fields on the host
// (crates/fastwarp-reg/src/grid.rs, volume.rs, field_ops.rs `VectorField`).
//
// A grid is (size, spacing, origin, direction) in LPS millimetres: p = origin + D diag(s) index.
// Volumes and fields are x-fastest; a field stores one LPS-mm displacement per voxel.
#pragma once

#include "common.hpp"

namespace fw {

struct Grid {
  std::array<size_t, 3> size{1, 1, 1};
  Vec3 spacing{1.0, 1.0, 1.0};
  Vec3 origin{0.0, 0.0, 0.0};
  Mat3 direction = kIdentity3;  // axis directions as columns: direction[row][axis]
  Mat3 index_from_offset = kIdentity3;

  Grid() = default;

  // grid.rs:46-88 `Grid::new`: rejects zero sizes, overflowing voxel counts, non-positive or
  // non-finite spacing, non-finite origin/direction and singular directions.
  Grid(std::array<size_t, 3> sz, Vec3 sp, Vec3 org, Mat3 dir) : size(sz), spacing(sp), origin(org), direction(dir) {
    for (size_t s : size)
      if (s == 0) invalid("grid", "zero size");
    if (size[0] > (size_t(1) << 20) || size[1] > (size_t(1) << 20) || size[2] > (size_t(1) << 20))
      invalid("grid", "voxel count overflows");
    for (double s : spacing)
      if (!std::isfinite(s) || s <= 0.0) invalid("grid", "spacing must be positive and finite");
    for (double o : origin)
      if (!std::isfinite(o)) invalid("grid", "origin and direction must be finite");
    for (auto& row : direction)
      for (double v : row)
        if (!std::isfinite(v)) invalid("grid", "origin and direction must be finite");
    Mat3 scaled = direction;
    for (auto& row : scaled)
      for (int c = 0; c < 3; ++c) row[c] *= spacing[c];
    if (!mat3_inverse(scaled, index_from_offset)) invalid("grid", "singular direction matrix");
  }

  size_t num_voxels() const { return size[0] * size[1] * size[2]; }
  size_t linear_index(size_t i, size_t j, size_t k) const { return i + size[0] * (j + size[1] * k); }
  std::array<size_t, 3> ijk(size_t l) const { return {l % size[0], (l / size[0]) % size[1], l / (size[0] * size[1])}; }

  // grid.rs:126-139
  Vec3 index_to_point(const Vec3& idx) const {
    const Vec3 scaled{idx[0] * spacing[0], idx[1] * spacing[1], idx[2] * spacing[2]};
    const Vec3 r = mat3_mul_vec(direction, scaled);
    return {origin[0] + r[0], origin[1] + r[1], origin[2] + r[2]};
  }
  // grid.rs:142-151
  Vec3 point_to_index(const Vec3& p) const {
    return mat3_mul_vec(index_from_offset, {p[0] - origin[0], p[1] - origin[1], p[2] - origin[2]});
  }
  // grid.rs:154-156: ITK IsInsideBuffer, each coordinate in [-0.5, size - 0.5).
  bool contains_continuous_index(const Vec3& c) const {
    for (int a = 0; a < 3; ++a)
      if (!(c[a] >= -0.5 && c[a] < static_cast<double>(size[a]) - 0.5)) return false;
    return true;
  }

  bool operator==(const Grid& o) const {
    return size == o.size && spacing == o.spacing && origin == o.origin && direction == o.direction;
  }
  bool operator!=(const Grid& o) const { return !(*this == o); }

  // grid.rs:231-239 `approx_eq`
  bool approx_eq(const Grid& o, double tol) const {
    if (size != o.size) return false;
    for (int a = 0; a < 3; ++a) {
      if (std::fabs(spacing[a] - o.spacing[a]) > tol || std::fabs(origin[a] - o.origin[a]) > tol) return false;
      for (int b = 0; b < 3; ++b)
        if (std::fabs(direction[a][b] - o.direction[a][b]) > tol) return false;
    }
    return true;
  }
};

template <typename T>
struct Volume {
  Grid grid;
  HostVec<T> data;
  Volume() = default;
  Volume(const Grid& g, HostVec<T> d) : grid(g), data(std::move(d)) { check(); }
  Volume(const Grid& g, const std::vector<T>& d) : grid(g), data(d.begin(), d.end()) { check(); }
  static Volume zeros(const Grid& g) { return Volume(g, HostVec<T>(g.num_voxels(), T{})); }

 private:
  void check() const {
    if (data.size() != grid.num_voxels()) invalid("volume data", "length differs from the voxel count");
  }
};

using Image = Volume<double>;

// A read-only view of a caller's f64 volume (no copy): the same `grid` / `data` surface as Image
// for the code that only reads its inputs.
struct ImageView {
  struct Values {
    const double* p = nullptr;
    size_t n = 0;
    size_t size() const { return n; }
    const double& operator[](size_t i) const { return p[i]; }
    const double* data() const { return p; }
    const double* begin() const { return p; }
    const double* end() const { return p + n; }
  };
  Grid grid;
  Values data;
  ImageView() = default;
  ImageView(const Grid& g, const double* values) : grid(g), data{values, g.num_voxels()} {}
  explicit ImageView(const Image& im) : grid(im.grid), data{im.data.data(), im.data.size()} {}
};

// A displacement field on a grid, (dx, dy, dz) LPS mm per voxel, x-fastest (field_ops.rs:58-63).
struct VectorField {
  Grid grid;
  std::vector<Vec3> data;
  VectorField() = default;
  VectorField(const Grid& g, std::vector<Vec3> d) : grid(g), data(std::move(d)) {
    if (data.size() != grid.num_voxels()) invalid("vector field data", "length differs from the voxel count");
  }
  static VectorField zeros(const Grid& g) { return VectorField(g, std::vector<Vec3>(g.num_voxels(), Vec3{0, 0, 0})); }

  // field_ops.rs:81-91 `from_fn`
  template <typename F>
  static VectorField from_fn(const Grid& g, F f) {
    std::vector<Vec3> d(g.num_voxels());
    parallel_for(d.size(), [&](size_t b, size_t e) {
      for (size_t l = b; l < e; ++l) {
        auto [i, j, k] = g.ijk(l);
        d[l] = f(g.index_to_point({double(i), double(j), double(k)}));
      }
    });
    return VectorField(g, std::move(d));
  }
};

}  // namespace fw
