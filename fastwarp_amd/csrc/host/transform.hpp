// Affine and parametric (Euler3D / affine) transforms with ITK MatrixOffsetTransformBase
// semantics and their parameter Jacobians (crates/fastwarp-reg/src/transform/affine.rs,
// transform/param.rs). Maps fixed (virtual) points to moving points, f64 throughout.
#pragma once

#include "common.hpp"

namespace fw {

// affine.rs:34-45: y = M (x - c) + c + t.
struct AffineTransform {
  Mat3 matrix = kIdentity3;
  Vec3 translation{0, 0, 0};
  Vec3 center{0, 0, 0};

  static AffineTransform identity() { return {}; }
  static AffineTransform from_translation(const Vec3& t) { return {kIdentity3, t, {0, 0, 0}}; }
  static AffineTransform from_offset(const Mat3& m, const Vec3& off) { return {m, off, {0, 0, 0}}; }

  // affine.rs:73-77: ITK's offset t + c - M c.
  Vec3 offset() const { return vsub(vadd(translation, center), mat3_mul_vec(matrix, center)); }
  // affine.rs:80-83
  Vec3 apply(const Vec3& p) const {
    const Vec3 r = mat3_mul_vec(matrix, vsub(p, center));
    return vadd(vadd(r, center), translation);
  }
  // affine.rs:86-91: same centre, translation -M^-1 t.
  AffineTransform inverse() const {
    Mat3 inv;
    if (!mat3_inverse(matrix, inv)) invalid("transform", "singular affine matrix");
    const Vec3 t = mat3_mul_vec(inv, translation);
    return {inv, {-t[0], -t[1], -t[2]}, center};
  }
  // affine.rs:94-98: outer(inner(p)) with a zero centre.
  static AffineTransform compose(const AffineTransform& outer, const AffineTransform& inner) {
    const Mat3 m = mat3_mul(outer.matrix, inner.matrix);
    const Vec3 off = vadd(mat3_mul_vec(outer.matrix, inner.offset()), outer.offset());
    return from_offset(m, off);
  }
  std::array<double, 12> parameters() const {
    std::array<double, 12> out{};
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) out[r * 3 + c] = matrix[r][c];
      out[9 + r] = translation[r];
    }
    return out;
  }
  // affine.rs:153-186: Euler3D matrix Rz Rx Ry (Rz Ry Rx with the ZYX flag).
  static AffineTransform from_euler3d(const double* p, const Vec3& center, bool zyx = false) {
    const double cx = std::cos(p[0]), sx = std::sin(p[0]);
    const double cy = std::cos(p[1]), sy = std::sin(p[1]);
    const double cz = std::cos(p[2]), sz = std::sin(p[2]);
    const Mat3 rx = {{{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}}};
    const Mat3 ry = {{{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}}};
    const Mat3 rz = {{{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}}};
    const Mat3 m = zyx ? mat3_mul(mat3_mul(rz, ry), rx) : mat3_mul(mat3_mul(rz, rx), ry);
    return {m, {p[3], p[4], p[5]}, center};
  }
  static AffineTransform from_parameters12(const double* p, const Vec3& center) {
    AffineTransform t;
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) t.matrix[r][c] = p[r * 3 + c];
      t.translation[r] = p[9 + r];
    }
    t.center = center;
    return t;
  }
};

// linear.rs:1071-1076 `with_center`: the same map expressed about another centre.
inline AffineTransform with_center(const AffineTransform& a, const Vec3& c) {
  const Vec3 off = a.offset();
  const Vec3 mc = mat3_mul_vec(a.matrix, c);
  return {a.matrix, vadd(vsub(off, c), mc), c};
}

enum class ParamKind { Euler3d, Affine };

// param.rs:36-52: the optimised transform of a linear stage.
struct ParametricTransform {
  ParamKind kind = ParamKind::Euler3d;
  std::array<double, 12> params{};
  Vec3 center{0, 0, 0};
  bool zyx = false;

  static ParametricTransform euler3d(const std::array<double, 6>& p, const Vec3& c) {
    ParametricTransform t;
    t.kind = ParamKind::Euler3d;
    for (int i = 0; i < 6; ++i) t.params[i] = p[i];
    t.center = c;
    return t;
  }
  static ParametricTransform affine(const AffineTransform& a) {
    ParametricTransform t;
    t.kind = ParamKind::Affine;
    t.params = a.parameters();
    t.center = a.center;
    return t;
  }
  size_t count() const { return kind == ParamKind::Euler3d ? 6 : 12; }
  std::vector<double> parameters() const { return std::vector<double>(params.begin(), params.begin() + count()); }
  void set_parameters(const std::vector<double>& v) {
    if (v.size() != count()) invalid("transform parameters", "wrong number of parameters");
    for (size_t i = 0; i < v.size(); ++i) params[i] = v[i];
  }
  AffineTransform to_affine() const {
    return kind == ParamKind::Euler3d ? AffineTransform::from_euler3d(params.data(), center, zyx)
                                      : AffineTransform::from_parameters12(params.data(), center);
  }

  // param.rs:218-300 `JacobianCtx::columns`: dT/dtheta_k at p, one column per parameter.
  size_t jacobian(const Vec3& p, std::array<Vec3, 12>& out) const {
    if (kind == ParamKind::Affine) {
      for (auto& c : out) c = {0, 0, 0};
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out[r * 3 + c][r] = p[c] - center[c];
        out[9 + r][r] = 1.0;
      }
      return 12;
    }
    const double cx = std::cos(params[0]), sx = std::sin(params[0]);
    const double cy = std::cos(params[1]), sy = std::sin(params[1]);
    const double cz = std::cos(params[2]), sz = std::sin(params[2]);
    const double px = p[0] - center[0], py = p[1] - center[1], pz = p[2] - center[2];
    for (int i = 0; i < 6; ++i) out[i] = {0, 0, 0};
    if (zyx) {
      out[0] = {(cz * sy * cx + sz * sx) * py + (-cz * sy * sx + sz * cx) * pz,
                (sz * sy * cx - cz * sx) * py + (-sz * sy * sx - cz * cx) * pz, (cy * cx) * py + (-cy * sx) * pz};
      out[1] = {(-cz * sy) * px + (cz * cy * sx) * py + (cz * cy * cx) * pz,
                (-sz * sy) * px + (sz * cy * sx) * py + (sz * cy * cx) * pz, (-cy) * px + (-sy * sx) * py + (-sy * cx) * pz};
      out[2] = {(-sz * cy) * px + (-sz * sy * sx - cz * cx) * py + (-sz * sy * cx + cz * sx) * pz,
                (cz * cy) * px + (cz * sy * sx - sz * cx) * py + (cz * sy * cx + sz * sx) * pz, 0.0};
    } else {
      out[0] = {(-sz * cx * sy) * px + (sz * sx) * py + (sz * cx * cy) * pz,
                (cz * cx * sy) * px + (-cz * sx) * py + (-cz * cx * cy) * pz, (sx * sy) * px + (cx)*py + (-sx * cy) * pz};
      out[1] = {(-cz * sy - sz * sx * cy) * px + (cz * cy - sz * sx * sy) * pz,
                (-sz * sy + cz * sx * cy) * px + (sz * cy + cz * sx * sy) * pz, (-cx * cy) * px + (-cx * sy) * pz};
      out[2] = {(-sz * cy - cz * sx * sy) * px + (-cz * cx) * py + (-sz * sy + cz * sx * cy) * pz,
                (cz * cy - sz * sx * sy) * px + (-sz * cx) * py + (cz * sy + sz * sx * cy) * pz, 0.0};
    }
    for (int d = 0; d < 3; ++d) out[3 + d][d] = 1.0;
    return 6;
  }
};

// param.rs:306-341: x -> outer(optimised(x)).
struct LinearMovingTransform {
  AffineTransform outer;
  ParametricTransform optimised;

  Vec3 apply(const Vec3& p) const { return outer.apply(optimised.to_affine().apply(p)); }
  LinearMovingTransform with_parameters(const std::vector<double>& v) const {
    LinearMovingTransform t = *this;
    t.optimised.set_parameters(v);
    return t;
  }
  // outer(optimised(x)) as one affine (zero centre), for device evaluation.
  AffineTransform collapsed() const { return AffineTransform::compose(outer, optimised.to_affine()); }
  // param.rs:357-364: optimised Jacobian columns left-multiplied by the outer matrix.
  size_t jacobian(const Vec3& p, std::array<Vec3, 12>& out) const {
    const size_t n = optimised.jacobian(p, out);
    for (size_t i = 0; i < n; ++i) out[i] = mat3_mul_vec(outer.matrix, out[i]);
    return n;
  }
};

}  // namespace fw
