// Checks of the product's host code (the C++ drivers' building blocks) against the reference's
// published values and closed forms. Built and run by tests/test_host.py; prints "ok <name>" per
// passing check and exits non-zero on the first failure.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <random>
#include <sstream>
#include <thread>

#include <unistd.h>

#include "../../fastwarp_amd/csrc/hip/fastdiv.hpp"
#include "../../fastwarp_amd/csrc/host/admission.hpp"
#include "../../fastwarp_amd/csrc/host/gauss_taps.hpp"
#include "../../fastwarp_amd/csrc/host/hdf5.hpp"
#include "../../fastwarp_amd/csrc/host/itk_h5.hpp"
#include "../../fastwarp_amd/csrc/host/linear.hpp"
#include "../../fastwarp_amd/csrc/host/mi_host.hpp"
#include "../../fastwarp_amd/csrc/host/preprocess.hpp"
#include "../../fastwarp_amd/csrc/host/request_loop.hpp"
#include "../../fastwarp_amd/csrc/host/tuning.hpp"
#include "../../oracle/oracle.hpp"

using namespace fw;

#define CHECK(cond, name)                                        \
  do {                                                           \
    if (!(cond)) {                                               \
      std::printf("FAIL %s (%s:%d)\n", name, __FILE__, __LINE__); \
      std::exit(1);                                              \
    }                                                            \
  } while (0)

static void ok(const char* n) { std::printf("ok %s\n", n); }

// sampling.rs:243-266: the published MT19937 outputs for seed 5489.
static void mt19937() {
  Mt19937 mt(5489);
  const uint32_t first[10] = {3499211612u, 581869302u, 3890346734u, 3586334585u, 545404204u,
                              4161255391u, 3922919429u, 949333985u,  2715962298u, 1323567403u};
  for (uint32_t v : first) CHECK(mt.integer() == v, "mt19937 first outputs");
  Mt19937 m2(5489);
  uint32_t last = 0;
  for (int i = 0; i < 10000; ++i) last = m2.integer();
  CHECK(last == 4123659995u, "mt19937 10000th output");
  // the array form of the stream (sample jitter draws) equals the generator call by call, across
  // many reloads, and gives the published 10000th output
  for (uint32_t seed : {5489u, 20260910u, 0u, 4294967295u}) {
    const HostVec<uint32_t> raw = mt19937_untempered(seed, 10000);
    Mt19937 g(seed);
    for (size_t i = 0; i < 10000; ++i) CHECK(mt19937_temper(raw[624 + i]) == g.integer(), "mt19937 array stream");
  }
  CHECK(mt19937_temper(mt19937_untempered(5489u, 10000)[624 + 9999]) == 4123659995u, "mt19937 array 10000th output");
  ok("mt19937");
}

// sampling.rs:302-340: REGULAR sampling takes every 4th voxel, jittered with the generator's
// Box-Muller draws in order (x first), and the seed advances once per call.
static void regular_sampling() {
  const Grid g({8, 4, 2}, {1, 1, 1}, {0, 0, 0}, kIdentity3);
  SeedSequence seeds(20260910u);
  auto a = metric_sample_points(g, 0.25f, seeds);
  CHECK(a.size() == 16, "regular sample count");
  CHECK(seeds.peek() == 20260911u, "seed advanced once");
  Mt19937 rng(20260910u);
  for (int q = 0; q < 2; ++q) {
    const Vec3 base = g.index_to_point({double((q * 4) % 8), double((q * 4 / 8) % 4), 0.0});
    for (int d = 0; d < 3; ++d) {
      const uint32_t u1 = rng.integer(), u2 = rng.integer();
      const double want = base[d] + normal_from(u1, u2, 0.0, 1.0) / 3.0;
      CHECK(std::fabs(a[q][d] - want) < 1e-15, "regular jitter");
    }
  }
  ok("regular_sampling");
}

// pyramid.rs:478-494: the ITK shrink rule on the 1 mm MNI-like grid and the ANTs factors.
static void shrink() {
  const Grid g({197, 233, 189}, {1, 1, 1}, {-98, -126, -72}, kIdentity3);
  const Grid s = shrink_grid(g, {8, 8, 8});
  CHECK(s.size[0] == 24 && s.size[1] == 29 && s.size[2] == 23, "shrink sizes");
  CHECK(std::fabs(s.origin[0] - (-98.0 + (98.0 - 11.5 * 8.0))) < 1e-12, "shrink origin x");
  CHECK(std::fabs(s.origin[1] - (-126.0 + (116.0 - 14.0 * 8.0))) < 1e-12, "shrink origin y");
  CHECK(std::fabs(s.origin[2] - (-72.0 + (94.0 - 11.0 * 8.0))) < 1e-12, "shrink origin z");
  auto f = shrink_factors_per_dimension(4, {1.0, 1.0, 2.0});
  CHECK(f[0] == 4 && f[1] == 4 && f[2] == 2, "factors 4 anisotropic");
  f = shrink_factors_per_dimension(4, {0.5, 0.5, 3.0});
  CHECK(f[0] == 4 && f[1] == 4 && f[2] == 1, "factors thick slices");
  f = shrink_factors_per_dimension(8, {2.0, 1.0, 1.0});
  CHECK(f[0] == 4 && f[1] == 8 && f[2] == 8, "factors smallest spacing");
  const Grid one = shrink_grid(Grid({7, 6, 5}, {1, 2, 3}, {1, 2, 3}, kIdentity3), {100, 100, 100});
  CHECK(one.size[0] == 1 && one.size[1] == 1 && one.size[2] == 1, "collapse to one voxel");
  ok("shrink");
}

// The HDF5 container (hdf5.rs:1465-1560 cases): lookup3's published answers; the ITK file written
// by libhdf5 1.12.1 (tests/golden/itk/ctest.h5) read with every checksum verified; a tree of
// groups, strings and f64 values written and read back bit for bit, with wide size fields for a
// long name and a large dataset; damaged files refused.
static void hdf5_container() {
  const std::string four = "Four score and seven years ago";
  CHECK(h5::lookup3(reinterpret_cast<const uint8_t*>(""), 0) == 0xdeadbeefu, "lookup3 empty");
  CHECK(h5::lookup3(reinterpret_cast<const uint8_t*>(four.data()), four.size()) == 0x17770551u, "lookup3 four score");
  std::string here = __FILE__;
  const std::string golden = here.substr(0, here.rfind('/')) + "/../golden/itk/ctest.h5";
  const h5::File f = h5::File::open(golden);
  const std::vector<std::string> root = f.members("");
  CHECK(std::find(root.begin(), root.end(), "TransformGroup") != root.end(), "ctest root group");
  CHECK(f.read_string("ITKVersion").rfind("5", 0) == 0, "ctest ITK version");
  CHECK((f.members("TransformGroup") == std::vector<std::string>{"0", "1", "2"}), "ctest members");
  CHECK(f.read_string("TransformGroup/0/TransformType") == "CompositeTransform_double_3_3", "ctest composite");
  CHECK(f.read_string("TransformGroup/1/TransformType") == "AffineTransform_double_3_3", "ctest affine");
  CHECK(f.read_doubles("TransformGroup/1/TransformParameters").size() == 12, "ctest affine parameters");
  CHECK(f.read_doubles("TransformGroup/2/TransformFixedParameters").size() == 18, "ctest field fixed parameters");
  CHECK(!f.exists("TransformGroup/3"), "ctest has three members");
  CHECK(f.is_group("TransformGroup/1") && !f.is_group("TransformGroup/1/TransformType"), "ctest groups");

  const std::vector<double> vals = {0.1, -0.0, std::numeric_limits<double>::denorm_min(), 1e300, -1234.5678901234567};
  std::vector<float> widen = {1.5f, -0.0f, std::numeric_limits<float>::denorm_min(), 3.0e38f};
  std::vector<double> big(5000);
  for (size_t i = 0; i < big.size(); ++i) big[i] = double(i) * 0.25 - 7.0;
  const std::string long_name(300, 'x');
  std::ostringstream out;
  h5::write(out, {h5::Node::string("Name", "hello"),
                  h5::Node::group("G", {h5::Node::values("v", vals), h5::Node::values("empty", {}),
                                        h5::Node::string("s", "world wide"),
                                        h5::Node::widened("w", widen.data(), widen.size())}),
                  h5::Node::values(long_name, big)});
  const std::string bytes = out.str();
  const h5::File g(std::vector<uint8_t>(bytes.begin(), bytes.end()));
  CHECK((g.members("") == std::vector<std::string>{"Name", "G", long_name}), "written members");
  CHECK(g.read_string("Name") == "hello" && g.read_string("G/s") == "world wide", "written strings");
  const std::vector<double> back = g.read_doubles("G/v");
  bool same = back.size() == vals.size();
  for (size_t i = 0; same && i < vals.size(); ++i) same = std::memcmp(&back[i], &vals[i], 8) == 0;
  CHECK(same, "written f64 values bit for bit");
  const std::vector<double> wb = g.read_doubles("G/w");
  same = wb.size() == widen.size();
  for (size_t i = 0; same && i < widen.size(); ++i) {
    const double d = double(widen[i]);
    same = std::memcmp(&wb[i], &d, 8) == 0;
  }
  CHECK(same, "widened f32 values");
  CHECK(g.read_doubles("G/empty").empty() && g.read_doubles(long_name) == big, "empty and large datasets");
  bool refused = false;
  try {
    g.read_doubles("Name");
  } catch (const Error&) {
    refused = true;
  }
  CHECK(refused, "a string is not a float dataset");
  // a flipped byte in any object header or in the superblock fails its checksum
  const auto fails = [](std::vector<uint8_t> b) {
    try {
      const h5::File x(std::move(b));
      x.members("");
      x.members("TransformGroup");
      for (const char* m : {"0", "1", "2"}) x.read_string(std::string("TransformGroup/") + m + "/TransformType");
      return false;
    } catch (const Error&) {
      return true;
    }
  };
  std::ifstream in(golden, std::ios::binary);
  const std::vector<uint8_t> ctest((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  CHECK(!fails(ctest), "ctest unchanged reads");
  for (size_t at : {size_t(20), size_t(60)}) {
    std::vector<uint8_t> bad = ctest;
    bad[at] ^= 0x40;
    CHECK(fails(bad), "a flipped byte is refused");
  }
  CHECK(fails(std::vector<uint8_t>(ctest.begin(), ctest.begin() + 40)), "a truncated file is refused");
  CHECK(fails(std::vector<uint8_t>(10, 0)), "not an HDF5 file");
  ok("hdf5_container");
}

// itk_h5.rs:205-260: the ITK file is [affine, field] as its values say; a chain written by the
// product reads back with the same bits; an unknown class and a non-numeric member are refused.
static void itk_composite() {
  std::string here = __FILE__;
  const std::string golden = here.substr(0, here.rfind('/')) + "/../golden/itk/ctest.h5";
  const std::vector<ChainElement> c = read_composite_h5(golden);
  CHECK(c.size() == 2 && !c[0].is_field && c[1].is_field, "ctest chain");
  CHECK(c[0].affine.matrix[0][0] == 2.0 && c[0].affine.matrix[1][1] == 1.0 && c[0].affine.matrix[2][2] == 1.0 &&
            c[0].affine.matrix[0][1] == 0.0 && c[0].affine.translation[0] == 0.0 && c[0].affine.center[2] == 0.0,
        "ctest affine values");
  CHECK(c[1].grid.size[0] == 6 && c[1].grid.size[1] == 5 && c[1].grid.size[2] == 4 && c[1].field.size() == 360,
        "ctest field grid");
  bool unit = true;
  for (size_t v = 0; v < 120; ++v) unit = unit && c[1].field[3 * v] == 1.0f && c[1].field[3 * v + 1] == 0.0f && c[1].field[3 * v + 2] == 0.0f;
  CHECK(unit, "ctest field values");

  const std::string tmp = "/tmp/fw_host_checks_" + std::to_string(::getpid()) + "_Composite.h5";
  std::vector<ChainElement> chain(2);
  chain[0].affine.matrix = {{{1.01, 0.02, 0.0}, {-0.02, 0.99, 0.01}, {0.0, 0.0, 1.0}}};
  chain[0].affine.translation = {1.0, -2.0, 0.5};
  chain[0].affine.center = {3.0, 4.0, 5.0};
  chain[1].is_field = true;
  chain[1].grid = Grid({5, 6, 7}, {2.0, 2.5, 3.0}, {-4.0, -5.0, -6.0}, {{{0.0, -1.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 0.0, 1.0}}});
  chain[1].field.resize(3 * 5 * 6 * 7);
  for (size_t i = 0; i < chain[1].field.size(); ++i) chain[1].field[i] = float(std::sin(0.37 * double(i)) * 3.0);
  write_composite_h5(tmp, chain);
  const std::vector<ChainElement> back = read_composite_h5(tmp);
  CHECK(back.size() == 2 && !back[0].is_field && back[1].is_field, "written chain kinds");
  CHECK(back[0].affine.parameters() == chain[0].affine.parameters() && back[0].affine.center == chain[0].affine.center,
        "written affine bits");
  CHECK(back[1].grid.size == chain[1].grid.size && back[1].grid.spacing == chain[1].grid.spacing &&
            back[1].grid.origin == chain[1].grid.origin && back[1].grid.direction == chain[1].grid.direction,
        "written field grid");
  CHECK(back[1].field == chain[1].field, "written field bits");
  const h5::File w = h5::File::open(tmp);
  CHECK((w.members("") == std::vector<std::string>{"HDFVersion", "ITKVersion", "OSName", "OSVersion", "TransformGroup"}),
        "written root layout");
  CHECK(w.read_string("TransformGroup/0/TransformType") == "CompositeTransform_double_3_3" &&
            w.read_string("TransformGroup/2/TransformType") == "DisplacementFieldTransform_double_3_3",
        "written type names");
  // the writer's dataset messages are the bytes libhdf5 1.12.1 wrote into the fixture for the same
  // kinds of dataset: dataspace, datatype (IEEE f64; variable-length string) and fill value bodies
  // equal, layout version 3 contiguous; for a string, a 12-value and an 18-value f64 dataset
  const h5::File f = h5::File::open(golden);
  const auto body = [](const h5::File& file, const std::string& path, uint8_t type) {
    for (const auto& m : file.header_messages(path))
      if (m.first == type) return m.second;
    return std::vector<uint8_t>{};
  };
  for (const char* ds : {"TransformGroup/1/TransformType", "TransformGroup/1/TransformParameters",
                         "TransformGroup/2/TransformFixedParameters"}) {
    for (uint8_t type : {uint8_t(0x01), uint8_t(0x03), uint8_t(0x05)})
      CHECK(!body(f, ds, type).empty() && body(f, ds, type) == body(w, ds, type), "message bodies as libhdf5 writes them");
    const std::vector<uint8_t> lf = body(f, ds, 0x08), lw = body(w, ds, 0x08);
    CHECK(lf.size() == 18 && lw.size() == 18 && lf[0] == 3 && lw[0] == 3 && lf[1] == 1 && lw[1] == 1 &&
              std::equal(lf.begin() + 10, lf.end(), lw.begin() + 10),
          "contiguous layout of the same size");
  }
  const std::vector<double> fixed = w.read_doubles("TransformGroup/2/TransformFixedParameters");
  CHECK(fixed.size() == 18 && fixed[0] == 5.0 && fixed[3] == -4.0 && fixed[6] == 2.0 && fixed[9 + 1] == -1.0,
        "field fixed parameters: size, origin, spacing, direction row-major");
  // an unknown class and a non-numeric member
  const auto refused = [&](std::vector<h5::Node> members) {
    h5::write_file(tmp, {h5::Node::group("TransformGroup", std::move(members))});
    try {
      read_composite_h5(tmp);
      return false;
    } catch (const Error&) {
      return true;
    }
  };
  // itk_params.rs:58-102: Euler3D (ITK's Rz Rx Ry, or Rz Ry Rx with the ZYX flag), translation and
  // float forms become affines
  {
    const std::vector<double> euler = {0.1, -0.2, 0.3, 1.0, 2.0, 3.0};
    h5::write_file(tmp, {h5::Node::group(
                       "TransformGroup",
                       {h5::Node::group("0", {h5::Node::string("TransformType", "CompositeTransform_double_3_3")}),
                        h5::Node::group("1", {h5::Node::string("TransformType", "Euler3DTransform_double_3_3"),
                                              h5::Node::values("TransformFixedParameters", {4.0, 5.0, 6.0}),
                                              h5::Node::values("TransformParameters", euler)}),
                        h5::Node::group("2", {h5::Node::string("TransformType", "TranslationTransform_float_3_3"),
                                              h5::Node::values("TransformFixedParameters", {}),
                                              h5::Node::values("TransformParameters", {0.5, -1.5, 2.5})}),
                        h5::Node::group("10", {h5::Node::string("TransformType", "Euler3DTransform_double_3_3"),
                                               h5::Node::values("TransformFixedParameters", {4.0, 5.0, 6.0, 1.0}),
                                               h5::Node::values("TransformParameters", euler)})})});
    const std::vector<ChainElement> e = read_composite_h5(tmp);
    CHECK(e.size() == 3 && !e[0].is_field && !e[1].is_field && !e[2].is_field, "euler and translation records");
    // Rz Rx Ry at (0.1, -0.2, 0.3): the rotation's (0, 2) entry is cz sy + sz sx cy
    const double cx = std::cos(0.1), sx = std::sin(0.1), cy = std::cos(-0.2), sy = std::sin(-0.2);
    const double cz = std::cos(0.3), sz = std::sin(0.3);
    CHECK(std::fabs(e[0].affine.matrix[0][2] - (cz * sy + sz * sx * cy)) < 1e-15 && e[0].affine.center[2] == 6.0 &&
              e[0].affine.translation[1] == 2.0,
          "euler3d matrix, centre and translation");
    CHECK(e[1].affine.translation[0] == 0.5 && e[1].affine.translation[2] == 2.5 && e[1].affine.matrix[1][1] == 1.0,
          "translation record");
    // member "10" sorts after "2" (numeric order); with the ZYX flag the matrix is Rz Ry Rx: (0, 2) = cz sy cx + sz sx
    CHECK(std::fabs(e[2].affine.matrix[0][2] - (cz * sy * cx + sz * sx)) < 1e-15, "euler3d zyx and numeric member order");
  }
  CHECK(refused({h5::Node::group("1", {h5::Node::string("TransformType", "BSplineTransform_double_3_3"),
                                       h5::Node::values("TransformFixedParameters", {0.0}),
                                       h5::Node::values("TransformParameters", {0.0})})}),
        "unknown class refused");
  CHECK(refused({h5::Node::group("x", {h5::Node::string("TransformType", "CompositeTransform_double_3_3")})}),
        "non-numeric member refused");
  CHECK(refused({h5::Node::group("0", {h5::Node::string("TransformType", "CompositeTransform_double_3_3")})}),
        "a file without a transform refused");
  std::remove(tmp.c_str());
  ok("itk_composite");
}

// pyramid.rs:544-584: unit DC gain, the requested variance, a preserved constant (the oracle's
// line filter); the product's coefficients, derived as polynomial products, against the oracle's
// restatement of the reference's closed forms.
static void deriche() {
  const Grid g({81, 4, 4}, {2.0, 1.0, 1.0}, {0, 0, 0}, kIdentity3);
  std::vector<double> d(g.num_voxels(), 0.0);
  d[g.linear_index(40, 1, 1)] = 1.0;
  HostVec<double> line(d.begin(), d.end());
  fw::oracle::deriche_axis(line, g.size, 0, fw::oracle::deriche_coefficients(6.0, 2.0));
  double sum = 0, var = 0;
  for (size_t i = 0; i < 81; ++i) {
    const double w = line[g.linear_index(i, 1, 1)];
    sum += w;
    var += w * (double(i) - 40.0) * (double(i) - 40.0);
  }
  CHECK(std::fabs(sum - 1.0) < 1e-6, "deriche dc gain");
  CHECK(std::fabs(var - 9.0) < 0.9, "deriche variance");
  // the Gaussian's shape to 1 % of its peak and a symmetric response to 1e-6 of the peak
  const double peak = 1.0 / (3.0 * std::sqrt(2.0 * 3.14159265358979323846));
  double worst = 0.0;
  for (size_t i = 0; i < 81; ++i) {
    const double x = double(i) - 40.0;
    worst = std::max(worst, std::fabs(line[g.linear_index(i, 1, 1)] - peak * std::exp(-x * x / 18.0)));
  }
  CHECK(worst < 0.01 * peak, "deriche shape");
  for (size_t i = 0; i < 40; ++i)
    CHECK(std::fabs(line[g.linear_index(40 - i, 1, 1)] - line[g.linear_index(40 + i, 1, 1)]) < 1e-6 * peak + 1e-12,
          "deriche symmetry");
  const Grid c({4, 4, 4}, {1, 1, 1}, {0, 0, 0}, kIdentity3);
  Image k(c, std::vector<double>(64, 7.0));
  Image s = fw::oracle::smooth_recursive_gaussian(k, {1.5, 1.5, 1.5});
  for (double v : s.data) CHECK(std::fabs(v - 7.0) < 1e-9, "deriche constant");
  // coefficients over sigma 0.3 .. 40 voxels. The reference's closed forms lose digits as the
  // poles approach 1: D(1) = 1 + d1 + d2 + d3 + d4 cancels terms up to cond = (1 + sum |d_k|) / |D(1)|
  // times larger, so the bound is 1e-14 of the largest coefficient times cond (cond is 1.0 at
  // sigma 0.3 voxels, 2.7e6 at 40). A dropped term, a wrong sign or a swapped pair of the fit
  // constants moves some coefficient by more than 1e-4 of the largest.
  for (double sd : {0.3, 0.5, 0.8, 1.0, 1.7, 2.5, 4.0, 6.5, 10.0, 17.0, 40.0}) {
    for (double spacing : {0.7, 1.0, 2.0}) {
      const Deriche p = Deriche::zero_order(sd * spacing, spacing);
      const fw::oracle::DericheCoefs o = fw::oracle::deriche_coefficients(sd * spacing, spacing);
      const double* pv[5] = {p.n, p.d, p.m, p.bn, p.bm};
      const double* ov[5] = {o.n, o.d, o.m, o.bn, o.bm};
      double scale = 0.0, abs_den = 1.0, den = 1.0;
      for (int q = 0; q < 4; ++q) {
        abs_den += std::fabs(o.d[q]);
        den += o.d[q];
      }
      const double cond = abs_den / std::fabs(den);
      for (int f = 0; f < 5; ++f)
        for (int q = 0; q < 4; ++q) scale = std::max(scale, std::fabs(ov[f][q]));
      for (int f = 0; f < 5; ++f)
        for (int q = 0; q < 4; ++q)
          CHECK(std::fabs(pv[f][q] - ov[f][q]) <= 1e-14 * cond * scale, "deriche coefficients");
      // unit DC gain of the two recursions together, from the product's own coefficients
      const double dc = (p.n[0] + p.n[1] + p.n[2] + p.n[3] + p.m[0] + p.m[1] + p.m[2] + p.m[3]) /
                        (1.0 + p.d[0] + p.d[1] + p.d[2] + p.d[3]);
      CHECK(std::fabs(dc - 1.0) <= 1e-14 * cond, "deriche coefficients dc gain");
    }
  }
  ok("deriche");
}

// device.rs:919-928: FASTWARP_GPU_WAIT_TIMEOUT_S takes positive finite seconds, else 300.
static void wait_timeout() {
  CHECK(wait_timeout_from(nullptr) == 300.0, "timeout default");
  CHECK(wait_timeout_from("2.5") == 2.5, "timeout seconds");
  CHECK(wait_timeout_from(" 7 ") == 7.0, "timeout with spaces");
  CHECK(wait_timeout_from("0") == 300.0, "timeout zero");
  CHECK(wait_timeout_from("-1") == 300.0, "timeout negative");
  CHECK(wait_timeout_from("inf") == 300.0, "timeout infinite");
  CHECK(wait_timeout_from("nan") == 300.0, "timeout nan");
  CHECK(wait_timeout_from("soon") == 300.0, "timeout text");
  CHECK(wait_timeout_from("1e-9") == 1e-9, "timeout tiny");
  ok("wait_timeout");
}

// The tuning table: size classes, class entries over plain ones over the fallback, the file format
// (the autotuner's notes and invalid values skipped, a malformed file rejected whole).
static void tuning() {
  CHECK(TuningTable::size_class(1) == 0 && TuningTable::size_class(2) == 1 && TuningTable::size_class(3) == 1, "class small");
  CHECK(TuningTable::size_class(182ull * 218 * 182) == 22 && TuningTable::size_class(1ull << 22) == 22, "class 1 mm");
  TuningTable t;
  CHECK(t.get("cc.tile_z", 1000, 16) == 16, "fallback");
  t.set("cc.tile_z", 8);
  t.set("cc.tile_z@9", 32);
  CHECK(t.get("cc.tile_z", 1000, 16) == 32 && t.get("cc.tile_z", 1023, 16) == 32, "class entry");
  CHECK(t.get("cc.tile_z", 1024, 16) == 8, "plain entry");
  t.set("cc.tile_z", -1);
  CHECK(t.get("cc.tile_z", 1024, 16) == 16 && t.size() == 1, "removed");
  CHECK(TuningTable::valid("mi.threads", 512) && !TuningTable::valid("mi.threads", 384), "mi threads");
  CHECK(TuningTable::valid("gauss.one_pass", 1) && !TuningTable::valid("gauss.one_pass", 2), "one pass");
  CHECK(TuningTable::valid("batch.graphs", 0) && !TuningTable::valid("batch.graphs", 2), "graphs");
  CHECK(TuningTable::valid("mi.fused_slots", 0) && !TuningTable::valid("mi.fused_slots", -1), "fused slots");
  CHECK(TuningTable::valid("mi.batch_blocks", 32) && !TuningTable::valid("mi.batch_blocks", 0), "batch blocks");
  CHECK(TuningTable::valid("deriche.lds", 1) && !TuningTable::valid("deriche.lds", 2), "deriche lds");
  CHECK(!TuningTable::valid("cc.tile_z", 0) && !TuningTable::valid("nonsense", 1), "unknown knob");
  const char* path = "/tmp/fw_tuning_check.json";
  {
    std::ofstream f(path);
    f << "{\n \"_command\": \"python -m x\", \"_seconds\": 1.5, \"_list\": [1, {\"a\": 2}],\n"
         " \"cc.tile_z@22\": 32, \"mi.threads@20\": 384, \"inv.max_blocks\": 4096, \"bogus@3\": 5\n}\n";
  }
  TuningTable u;
  CHECK(u.load(path), "load");
  CHECK(u.size() == 2, "entries kept");
  CHECK(u.get("cc.tile_z", 182ull * 218 * 182, 16) == 32 && u.get("mi.threads", 1u << 20, 0) == 0, "loaded values");
  CHECK(u.get("inv.max_blocks", 5, 2048) == 4096, "loaded plain");
  {
    std::ofstream f(path);
    f << "[1, 2]";
  }
  CHECK(!u.load(path) && u.size() == 2, "malformed kept the old table");
  CHECK(!u.load("/nonexistent/tuning.json"), "missing file");
  std::remove(path);
  ok("tuning");
}

// convergence.rs:188-213: the closed form of the two-control-point fit.
static void convergence() {
  WindowConvergenceMonitor m(10);
  const double vals[10] = {9.0, 8.5, 8.2, 8.0, 7.9, 7.85, 7.82, 7.80, 7.79, 7.785};
  double total = 0;
  for (double v : vals) {
    m.add_energy_value(v);
    total += v;
  }
  const double eps = 1e-4;
  double d0 = 0, d1 = 0, o0 = 0, o1 = 0;
  for (int n = 0; n < 10; ++n) {
    double p = n / 9.0;
    if (std::fabs(p - 1.0) <= eps) p = 1.0 - eps;
    const double t0 = 1.0 - p, t1 = p, s = t0 * t0 + t1 * t1, y = vals[n] / total;
    o0 += t0 * t0;
    o1 += t1 * t1;
    d0 += y * t0 * t0 * t0 / s;
    d1 += y * t1 * t1 * t1 / s;
  }
  CHECK(std::fabs(m.convergence_value() - (d0 / o0 - d1 / o1)) < 1e-14, "monitor closed form");
  WindowConvergenceMonitor early(10);
  early.add_energy_value(1.0);
  CHECK(early.convergence_value() == std::numeric_limits<double>::max(), "monitor window not full");
  ok("convergence");
}

// optimizer.rs:526-548: the conjugate gradient with the golden section minimises a quadratic.
struct Quadratic {
  std::array<double, 6> target{0, 0, 0, 3, -2, 1};
  double value(const LinearMovingTransform& t) const {
    double s = 0;
    for (int i = 0; i < 6; ++i) s += (t.optimised.params[i] - target[i]) * (t.optimised.params[i] - target[i]);
    return s;
  }
  std::pair<double, std::vector<double>> value_and_derivative(const LinearMovingTransform& t) const {
    std::vector<double> d(6);
    for (int i = 0; i < 6; ++i) d[i] = -2.0 * (t.optimised.params[i] - target[i]);
    return {value(t), d};
  }
};

static void optimizer() {
  Quadratic q;
  std::array<Vec3, 8> corners{};
  for (int i = 0; i < 8; ++i) corners[i] = {(i & 1) ? 20.0 : -20.0, (i & 2) ? 20.0 : -20.0, (i & 4) ? 20.0 : -20.0};
  LinearMovingTransform t{AffineTransform::identity(), ParametricTransform::euler3d({0, 0, 0, 0, 0, 0}, {0, 0, 0})};
  CgSettings st;
  st.iterations = 200;
  st.convergence_threshold = 1e-12;
  auto [best, trace] = optimise_level(q, corners, t, st);
  CHECK(std::fabs(best.optimised.params[3] - 3.0) < 0.05, "cg x");
  CHECK(std::fabs(best.optimised.params[4] + 2.0) < 0.05, "cg y");
  CHECK(std::fabs(best.optimised.params[5] - 1.0) < 0.05, "cg z");
  CHECK(trace.best_value < 0.01 && trace.line_search_evaluations > 0, "cg trace");
  // optimizer.rs:477-510: scales are squared corner shifts per unit parameter
  std::array<Vec3, 8> cube{};
  for (int i = 0; i < 8; ++i) cube[i] = {(i & 1) ? 50.0 : -50.0, (i & 2) ? 50.0 : -50.0, (i & 4) ? 50.0 : -50.0};
  auto sc = estimate_scales(cube, t, 0.01);
  for (int i = 3; i < 6; ++i) CHECK(std::fabs(sc[i] - 1.0) < 1e-9, "translation scales");
  const double want = 2 * 50.0 * 50.0;
  for (int i = 0; i < 3; ++i) CHECK(std::fabs(sc[i] - want) / want < 1e-3, "rotation scales");
  CHECK(std::fabs(estimate_step_scale(cube, t, {0, 0, 0, 3, 4, 0}, 0.01) - 5.0) < 1e-9, "step scale");
  ok("optimizer");
}

// The speculative line search (two probes per pass) against the plain golden search it stands for:
// the same step and the same evaluation count on rugged 1D profiles, flat stretches, ties and the
// f64::MAX sentinel, for many brackets.
struct Profile {
  int kind;
  uint64_t evaluations = 0;
  double f(double s) const {
    switch (kind % 5) {
      case 0: return (s - 0.37) * (s - 0.37);
      case 1: return std::floor(8.0 * std::fabs(s - 0.8)) / 8.0;  // plateaus: ties in both directions
      case 2: return s > 1.3 ? std::numeric_limits<double>::max() : std::sin(9.0 * s) + 0.1 * s;
      case 3: return std::fabs(s - 0.01) < 0.3 ? std::numeric_limits<double>::max() : -s;
      default: return std::cos(23.0 * s) * std::exp(-s);
    }
  }
  double step_of(const LinearMovingTransform& t) const { return t.optimised.params[3]; }
  double value(const LinearMovingTransform& t) {
    ++evaluations;
    return f(step_of(t));
  }
  bool batch_pays() const { return true; }
  std::vector<double> values(const std::vector<LinearMovingTransform>& ts) {
    std::vector<double> v;
    for (const auto& t : ts) v.push_back(f(step_of(t)));
    return v;
  }
};

static void speculative_search() {
  const LinearMovingTransform t{AffineTransform::identity(), ParametricTransform::euler3d({0, 0, 0, 0, 0, 0}, {0, 0, 0})};
  const std::vector<double> dir{0, 0, 0, 1, 0, 0};
  int cases = 0;
  for (int kind = 0; kind < 5; ++kind)
    for (double lr : {0.05, 0.3, 0.5, 1.0, 1.7, 4.0})
      for (double eps : {0.2, 0.05, 0.001})
        for (uint32_t cap : {2u, 20u}) {
          CgSettings st;
          st.epsilon = eps;
          st.max_line_search_iterations = cap;
          Profile plain{kind}, spec{kind};
          uint64_t e_plain = 0, e_spec = 0;
          const double a = golden_search(plain, t, dir, st, 0.0, lr, 2.0 * lr, e_plain);
          const double b = golden_search_speculative(spec, t, dir, st, 0.0, lr, 2.0 * lr, e_spec);
          CHECK(a == b, "speculative search step");
          CHECK(e_plain == e_spec && e_plain == plain.evaluations, "speculative search evaluation count");
          ++cases;
        }
  CHECK(cases == 180, "speculative search cases");
  ok("speculative_search");
}

// transform/param.rs:398-430: the Euler and affine Jacobians against central differences.
static void jacobians() {
  for (int kind = 0; kind < 2; ++kind) {
    ParametricTransform t = kind == 0 ? ParametricTransform::euler3d({0.1, -0.2, 0.3, 1, 2, 3}, {5, -4, 2})
                                      : ParametricTransform::affine(AffineTransform{{{{1.02, 0.03, -0.01}, {-0.02, 0.97, 0.04}, {0.01, 0.02, 1.01}}}, {0.5, -0.3, 0.2}, {5, -4, 2}});
    const Vec3 p{12.0, -7.0, 31.0};
    std::array<Vec3, 12> cols;
    const size_t n = t.jacobian(p, cols);
    for (size_t k = 0; k < n; ++k) {
      auto up = t, dn = t;
      up.params[k] += 1e-6;
      dn.params[k] -= 1e-6;
      const Vec3 a = up.to_affine().apply(p), b = dn.to_affine().apply(p);
      for (int d = 0; d < 3; ++d) CHECK(std::fabs((a[d] - b[d]) / 2e-6 - cols[k][d]) < 1e-6, "jacobian column");
    }
  }
  ok("jacobians");
}

// The f32 thresholds reproduce the f64 fixed bin of every f32 intensity (wgpu_backend/mi/dense.rs
// idea; checked exhaustively near each threshold and on random values).
static void thresholds() {
  const MiBinning b = MiBinning::from_ranges(32, 0.0312, 0.9471, -0.2, 1.3);
  const std::vector<float> t = b.fixed_thresholds();
  CHECK(t.size() == 27, "threshold count");
  auto bin_of = [&](float v) {
    int k = 2;
    for (float th : t) k += v >= th ? 1 : 0;
    return size_t(k);
  };
  std::mt19937 rng(3);
  std::uniform_real_distribution<float> u(-0.2f, 1.2f);
  for (int i = 0; i < 200000; ++i) {
    const float v = u(rng);
    CHECK(bin_of(v) == b.fixed_bin(double(v)), "random bins");
  }
  for (float th : t) {
    CHECK(b.fixed_bin(double(th)) > b.fixed_bin(double(std::nextafter(th, -1.0f))), "threshold is a bin edge");
    CHECK(bin_of(th) == b.fixed_bin(double(th)) && bin_of(std::nextafter(th, -1.0f)) == b.fixed_bin(double(std::nextafter(th, -1.0f))),
          "edge bins");
  }
  ok("thresholds");
}

// The product's Gaussian taps equal the oracle's restatement.
static void taps() {
  for (double v : {0.25, 0.5, 1.0, 3.0, 8.0})
    for (size_t w : {3u, 21u, 1000u}) {
      const auto a = gaussian_operator_taps(v, 0.001, w);
      const auto b = oracle::gaussian_taps(v, 0.001, w);
      CHECK(a.size() == b.size(), "taps size");
      for (size_t i = 0; i < a.size(); ++i) CHECK(a[i] == b[i], "taps value");
    }
  ok("taps");
}

// preprocess.rs: the window maps the quantile range onto [0, 1]; quantiles (0, 1) of a ramp give
// its minimum and (within the histogram margin) its maximum.
static void preprocess() {
  const Grid g({100, 1, 1}, {1, 1, 1}, {0, 0, 0}, kIdentity3);
  std::vector<double> d(100);
  for (int i = 0; i < 100; ++i) d[i] = 3.0 + 0.5 * i;
  const Image v(g, d);
  const IntensityHistogram h(v.data);
  CHECK(std::fabs(h.quantile(0.0) - 3.0) < 1e-9, "quantile 0 is the minimum");
  CHECK(std::fabs(h.quantile(1.0) - (52.5 + 49.5 / 2560.0)) < 1e-4, "quantile 1 is the margined maximum");
  const Image p = preprocess_volume(v, 0.0f, 1.0f);
  CHECK(p.data[0] == 0.0 && p.data[99] <= 1.0 && p.data[99] > 0.99, "window to [0, 1]");
  const Image w = preprocess_volume(v, 0.1f, 0.9f);
  CHECK(w.data[0] == 0.0 && w.data[99] == 1.0, "winsorised ends clamp");
  // preprocess.rs:271-303: quantiles interpolate inside the bin like ITK's (a ramp of 2560 values
  // has its quartiles at 640 and 1920), and the result is invariant to affine intensity changes
  {
    std::vector<double> ramp(2560);
    for (size_t i = 0; i < ramp.size(); ++i) ramp[i] = double(i);
    const IntensityHistogram hq(ramp);
    CHECK(std::fabs(hq.quantile(0.25) - 640.0) < 5.0, "ramp first quartile");
    CHECK(std::fabs(hq.quantile(0.75) - 1920.0) < 5.0, "ramp third quartile");
    const Grid g5({500, 1, 1}, {1, 1, 1}, {0, 0, 0}, kIdentity3);
    std::vector<double> a(500), b(500);
    for (int i = 0; i < 500; ++i) {
      a[i] = double((i * 37) % 101);
      b[i] = 3.0 * a[i] + 40.0;
    }
    const Image pa = preprocess_volume(Image(g5, a), 0.005f, 0.995f);
    const Image pb = preprocess_volume(Image(g5, b), 0.005f, 0.995f);
    for (int i = 0; i < 500; ++i) CHECK(std::fabs(pa.data[i] - pb.data[i]) < 1e-5, "affine intensity invariance");
  }
  // the device's f32 bin rule (count of f32 thresholds at or below v) equals index_of on f32
  // intensities: random values over and beyond the range, every threshold and its neighbours
  {
    Mt19937 rt(99);
    std::vector<double> fv(30000);
    for (auto& x : fv) x = double(float(-0.3 + 1.9 * (rt.integer() / 4294967296.0)));
    const IntensityHistogram hf(fv);
    const std::vector<float> th = hf.f32_thresholds();
    CHECK(th.size() == 255, "255 thresholds");
    std::vector<float> probes;
    for (double x : fv) probes.push_back(float(x));
    for (float t : th) probes.insert(probes.end(), {t, std::nextafter(t, -1e30f), std::nextafter(t, 1e30f)});
    probes.insert(probes.end(), {-1e30f, 1e30f, float(hf.bin_maxima().back())});
    for (float x : probes) {
      size_t count = 0;
      for (float t : th) count += t <= x ? 1 : 0;
      CHECK(count == hf.index_of(double(x)), "f32 threshold bins");
    }
    // the histogram rebuilt from its range and counts has the same quantiles
    std::vector<uint64_t> counts(256, 0);
    for (double x : fv) ++counts[hf.index_of(x)];
    double lo = fv[0], hi = fv[0];
    for (double x : fv) lo = std::min(lo, x), hi = std::max(hi, x);
    const IntensityHistogram hc(lo, hi, counts);
    for (double q : {0.005, 0.2, 0.5, 0.995}) CHECK(hc.quantile(q) == hf.quantile(q), "quantiles from counts");
  }
  // the bin lookup equals ITK's search (the largest bin whose minimum is <= v, clamped at both
  // ends) on random values, on every bin edge and on the neighbours of each edge
  std::vector<double> r(20000);
  Mt19937 mt(7);
  for (auto& x : r) x = -2.0 + 7.0 * (mt.integer() / 4294967296.0);
  const IntensityHistogram hr(r);
  const auto& mn = hr.bin_minima();
  const auto& mx = hr.bin_maxima();
  std::vector<double> probes = r;
  for (double e : mn) probes.insert(probes.end(), {e, std::nextafter(e, -1e9), std::nextafter(e, 1e9)});
  probes.insert(probes.end(), {mx.back(), std::nextafter(mx.back(), -1e9), -1e9, 1e9});
  for (double v : probes) {
    size_t want;
    if (v < mn[0]) want = 0;
    else if (v >= mx.back()) want = mn.size() - 1;
    else {
      want = 0;
      for (size_t j = 0; j < mn.size(); ++j)
        if (mn[j] <= v) want = j;
    }
    CHECK(hr.index_of(v) == want, "histogram bin lookup");
  }
  ok("preprocess");
}

// The product's Mattes finish (which skips the Kahan steps that cannot change the sum) equals the
// oracle's restatement of mi.rs:702-769 bit for bit: value and derivative weights, on sparse and
// dense joint histograms of fixed-point counts with wide dynamic ranges.
static void mi_finish() {
  std::mt19937_64 rng(77);
  int cases = 0;
  for (size_t bins : {5u, 12u, 32u, 64u})
    for (double density : {0.02, 0.1, 0.3, 1.0})
      for (int rep = 0; rep < 25; ++rep) {
        std::vector<double> joint(bins * bins, 0.0), marg(bins, 0.0);
        uint64_t valid = 0;
        for (size_t f = 0; f < bins; ++f)
          for (size_t m = 0; m < bins; ++m) {
            if (double(rng() % 1000000) / 1e6 >= density) continue;
            // fixed-point counts (2^-32 units) from a few units up to 2^31 samples' worth: totals
            // beyond 2^53 units make the Kahan compensation non-zero
            const int shift = int(rng() % 62);
            const uint64_t c = (rng() >> 1) >> shift;
            joint[f * bins + m] = double(c) / 4294967296.0;
            const uint64_t n = 1 + c / 4294967296u;
            marg[f] += double(n);
            valid += n;
          }
        for (bool ratio : {false, true}) {
          bool fw_ok = true, or_ok = true;
          HistogramFinish a;
          oracle::HistogramResult b;
          try {
            a = finish_histogram(bins, 0.031, valid, joint, marg, ratio);
          } catch (const Error&) {
            fw_ok = false;
          }
          try {
            b = oracle::finish_histogram(bins, 0.031, valid, joint, marg, ratio);
          } catch (const Error&) {
            or_ok = false;
          }
          CHECK(fw_ok == or_ok, "mattes finish: the same inputs are rejected");
          if (!fw_ok) continue;
          CHECK(std::memcmp(&a.value, &b.value, sizeof(double)) == 0, "mattes finish: value bit for bit");
          CHECK(a.ratio.size() == b.ratio.size() &&
                    (a.ratio.empty() || std::memcmp(a.ratio.data(), b.ratio.data(), a.ratio.size() * sizeof(double)) == 0),
                "mattes finish: derivative weights bit for bit");
          ++cases;
        }
      }
  CHECK(cases > 600, "mattes finish: enough non-degenerate cases");
  ok("mi_finish");
}

// The server's request loop with stand-in slots: results and errors reach the caller that waits,
// no more requests run at once than slots and budget allow, they start in arrival order, waiting
// and running requests cancel, and a loop with requests in flight shuts down promptly.
static void request_loop() {
  struct Slot {
    uint32_t index;
    uint64_t pool = 0;  // idle bytes the slot keeps between requests
  };
  struct Job {
    int input = 0, output = 0, sleep_us = 0;
    int fail = 0;  // 1: invalid input, 2: a device error
  };
  auto none = [](Slot&) { return uint64_t(0); };
  auto keep = [](Slot&) {};
  std::atomic<int> running{0}, peak{0}, order{0};
  std::vector<int> starts(400, -1);
  auto run = [&](Slot&, Job& j, const volatile int32_t* cancel) {
    const int now = ++running;
    int p = peak.load();
    while (now > p && !peak.compare_exchange_weak(p, now)) {
    }
    starts[size_t(order++)] = j.input;
    for (int t = 0; t < j.sleep_us; t += 200) {
      if (__atomic_load_n(cancel, __ATOMIC_RELAXED)) {
        --running;
        throw Cancelled();
      }
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    --running;
    if (j.fail == 1) throw Error("test", "bad input " + std::to_string(j.input));
    if (j.fail == 2) throw Error("hip", "device lost");
    j.output = 2 * j.input + 1;
  };
  auto make = [](uint32_t i) { return std::make_unique<Slot>(Slot{i}); };
  {
    RequestLoop<Slot, Job> loop(3, 1000, make, run, none, keep);
    std::vector<uint64_t> ids;
    for (int i = 0; i < 40; ++i) ids.push_back(loop.submit(100, Job{i, 0, 1000, 0}));
    for (int i = 0; i < 40; ++i) {
      auto [o, j] = loop.wait(ids[size_t(i)]);
      CHECK(o.status == Ended::Ok && j.output == 2 * i + 1 && o.slot >= 0 && o.slot < 3, "request loop: results");
      CHECK(o.footprint == 100 && std::fabs(o.total_s - o.queued_s - o.run_s) < 1e-6, "request loop: timing");
    }
    CHECK(peak.load() == 3, "request loop: as many at once as slots");
    bool threw = false;
    try {
      loop.wait(ids[0]);
    } catch (const Error&) {
      threw = true;
    }
    CHECK(threw, "request loop: a ticket is waited for once");
  }
  {
    // one slot runs the requests one after another, in arrival order (with several, the order in
    // which running requests reach their first line is up to the threads)
    const int first = order.load();
    RequestLoop<Slot, Job> loop(1, 1000, make, run, none, keep);
    std::vector<uint64_t> ids;
    for (int i = 0; i < 20; ++i) ids.push_back(loop.submit(100, Job{100 + i, 0, 100, 0}));
    for (uint64_t id : ids) loop.wait(id);
    for (int i = 0; i < 20; ++i) CHECK(starts[size_t(first + i)] == 100 + i, "request loop: arrival order");
  }
  {
    peak = 0;
    RequestLoop<Slot, Job> loop(3, 250, make, run, none, keep);
    std::vector<uint64_t> ids;
    for (int i = 0; i < 12; ++i) ids.push_back(loop.submit(100, Job{i, 0, 1000, 0}));
    for (uint64_t id : ids) loop.wait(id);
    CHECK(peak.load() == 2, "request loop: the budget bounds what runs at once");
    bool threw = false;
    try {
      loop.submit(251, Job{});
    } catch (const Error&) {
      threw = true;
    }
    CHECK(threw, "request loop: a request beyond the budget is refused");
  }
  {
    RequestLoop<Slot, Job> loop(1, 1000, make, run, none, keep);
    const uint64_t a = loop.submit(10, Job{1, 0, 2000000, 0});
    const uint64_t b = loop.submit(10, Job{2, 0, 1000, 0});
    const uint64_t c = loop.submit(10, Job{3, 0, 1000, 1});
    const uint64_t d = loop.submit(10, Job{4, 0, 1000, 2});
    loop.cancel(b);
    auto [ob, jb] = loop.wait(b);
    CHECK(ob.status == Ended::Cancelled && ob.slot == -1 && jb.output == 0, "request loop: a waiting request cancels at once");
    while (running.load() == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
    loop.cancel(a);
    auto [oa, ja] = loop.wait(a);
    CHECK(oa.status == Ended::Cancelled && oa.slot == 0 && oa.run_s < 1.0, "request loop: a running request cancels");
    auto [oc, jc] = loop.wait(c);
    CHECK(oc.status == Ended::Invalid && oc.error.find("bad input 3") != std::string::npos, "request loop: an error reaches its caller");
    auto [od, jd] = loop.wait(d);
    CHECK(od.status == Ended::Device, "request loop: a device error is one");
    uint32_t w = 9, r = 9;
    uint64_t bytes = 9;
    loop.counts(&w, &r, &bytes);
    CHECK(w == 0 && r == 0 && bytes == 0, "request loop: empty at the end");
  }
  {
    const auto t0 = std::chrono::steady_clock::now();
    {
      RequestLoop<Slot, Job> loop(2, 1000, make, run, none, keep);
      for (int i = 0; i < 6; ++i) loop.submit(10, Job{i, 0, 2000000, 0});
      std::this_thread::sleep_for(std::chrono::milliseconds(5));
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    CHECK(s < 1.0, "request loop: shutdown cancels what is in flight");
  }
  bool threw = false;
  try {
    RequestLoop<Slot, Job> loop(
        3, 1000,
        [](uint32_t i) -> std::unique_ptr<Slot> {
          if (i == 1) throw Error("hip", "no device");
          return std::make_unique<Slot>(Slot{i});
        },
        run, none, keep);
  } catch (const Error& e) {
    threw = e.what_ == "hip";
  }
  CHECK(threw, "request loop: a slot that cannot start fails the loop");
  {
    // two callers wait for one ticket at once: one gets the request, the other an unknown ticket
    RequestLoop<Slot, Job> loop(1, 1000, make, run, none, keep);
    const uint64_t id = loop.submit(10, Job{7, 0, 20000, 0});
    std::atomic<int> got{0}, unknown{0};
    std::vector<std::thread> waiters;
    for (int t = 0; t < 2; ++t)
      waiters.emplace_back([&] {
        try {
          auto [o, j] = loop.wait(id);
          if (o.status == Ended::Ok && j.output == 15) ++got;
        } catch (const Error& e) {
          if (e.message.find("unknown ticket") != std::string::npos) ++unknown;
        }
      });
    for (auto& w : waiters) w.join();
    CHECK(got.load() == 1 && unknown.load() == 1, "request loop: a second concurrent wait is refused");
  }
  {
    // slots that keep their buffers between requests: what runs plus what the idle pools keep
    // never exceeds the budget, since a running request may not reuse its slot's kept buffers
    // (each run allocates its whole footprint beside them), and a head that does not fit makes
    // idle slots give their buffers back. Failing requests keep their buffers too.
    std::mutex mu;
    uint64_t device = 0, device_peak = 0;
    std::atomic<int> over{0};
    auto run_keep = [&](Slot& s, Job& j, const volatile int32_t*) {
      const uint64_t fp = uint64_t(j.input);
      {
        std::lock_guard<std::mutex> lk(mu);
        device += fp;
        device_peak = std::max(device_peak, device);
        if (device > 1000) ++over;
      }
      std::this_thread::sleep_for(std::chrono::microseconds(j.sleep_us));
      std::lock_guard<std::mutex> lk(mu);
      // the request's buffers stay in the pool: it keeps the larger of the old pool and this one
      const uint64_t kept = std::max(s.pool, fp);
      device = device - fp - s.pool + kept;
      s.pool = kept;
      if (j.fail) throw Error("test", "failed with its buffers kept");
    };
    auto held = [&](Slot& s) {
      std::lock_guard<std::mutex> lk(mu);
      return s.pool;
    };
    auto release = [&](Slot& s) {
      std::lock_guard<std::mutex> lk(mu);
      device -= s.pool;
      s.pool = 0;
    };
    std::mt19937 rng(3);
    {
      RequestLoop<Slot, Job> loop(4, 1000, make, run_keep, held, release);
      std::vector<uint64_t> ids;
      for (int i = 0; i < 300; ++i) {
        const int fp = 50 + int(rng() % 500);
        ids.push_back(loop.submit(uint64_t(fp), Job{fp, 0, int(rng() % 300), int(rng() % 7 == 0)}));
      }
      int ok_count = 0;
      for (uint64_t id : ids) ok_count += loop.wait(id).first.status == Ended::Ok;
      CHECK(ok_count > 200, "request loop: kept buffers: requests run");
    }
    CHECK(over.load() == 0 && device_peak <= 1000, "request loop: kept buffers count against the budget");
  }
  ok("request_loop");
}

// The inversion kernels' multiply-high division equals integer division for every divisor a grid
// axis can have and every index below 2^31: around each multiple of the divisor, at the top of the
// range, and at random.
static void fastdiv() {
  std::mt19937 rng(5);
  const uint32_t top = (uint32_t(1) << 31) - 1;
  for (uint32_t d = 1; d <= 4096; ++d) {
    const FastDiv f = make_fastdiv(d);
    auto check = [&](uint32_t n) { return fast_div(n, f) == n / d; };
    bool good = check(0) && check(top) && check(top - 1) && check(top - d);
    for (uint32_t k = 1; k < 64 && good; ++k) {
      const uint64_t q = uint64_t(k) * d;
      good = check(uint32_t(q - 1)) && check(uint32_t(q)) && check(uint32_t(q + 1));
    }
    for (uint64_t k = top / d; k > top / d - 4 && good; --k) good = check(uint32_t(k * d)) && check(uint32_t(k * d - 1));
    for (int r = 0; r < 2000 && good; ++r) good = check(rng() & top);
    CHECK(good, "fast division by a grid axis");
  }
  for (uint32_t d : {5000u, 65535u, 100003u, (uint32_t(1) << 20) + 7u, top})
    for (int r = 0; r < 20000; ++r) {
      const uint32_t n = rng() & top;
      CHECK(fast_div(n, make_fastdiv(d)) == n / d, "fast division by a large divisor");
    }
  ok("fastdiv");
}

// Admission by footprint: arrival order, the budget never exceeded, a head that does not fit
// holds the queue back until memory frees, cancelled waiters leave without a trace, and a
// simulated stream of random requests drains with every request run exactly once.
static void admission() {
  AdmissionQueue q(100);
  CHECK(q.fits_alone(100) && !q.fits_alone(101), "admission: a request larger than the budget never fits");
  CHECK(!q.can_admit(), "admission: nothing to admit in an empty queue");
  q.push(1, 40);
  q.push(2, 50);
  q.push(3, 30);
  q.push(4, 5);
  CHECK(q.can_admit() && q.admit() == 1, "admission: the head starts first");
  CHECK(q.can_admit() && q.admit() == 2, "admission: the next fits beside it");
  CHECK(q.in_use() == 90 && !q.can_admit(), "admission: a head that does not fit waits");
  CHECK(q.waiting() == 2 && q.running() == 2, "admission: counts");
  q.finish(1);
  CHECK(q.in_use() == 50 && q.can_admit() && q.admit() == 3, "admission: a finish frees its footprint");
  CHECK(q.remove(4) && !q.remove(4) && q.waiting() == 0, "admission: a waiting request can be removed once");
  q.finish(2);
  q.finish(3);
  q.finish(3);  // a second finish is ignored
  CHECK(q.in_use() == 0 && q.running() == 0, "admission: all memory back");
  // random stream over a few slots: arrivals, completions in random order
  std::mt19937 rng(11);
  AdmissionQueue r(1000);
  std::vector<uint64_t> running, order;
  uint64_t next = 1, started = 0;
  const size_t slots = 3;
  std::vector<uint64_t> size_of(401, 0);
  while (started < 400) {
    if (next <= 400 && rng() % 3 != 0) {
      size_of[next] = 50 + rng() % 600;
      r.push(next, size_of[next]);
      ++next;
    }
    while (running.size() < slots && r.can_admit()) {
      const uint64_t id = r.admit();
      order.push_back(id);
      running.push_back(id);
      ++started;
    }
    uint64_t sum = 0;
    for (uint64_t id : running) sum += size_of[id];
    CHECK(sum == r.in_use() && sum <= 1000, "admission: the budget holds at every step");
    if (!running.empty() && (rng() % 2 == 0 || next > 400)) {
      const size_t k = rng() % running.size();
      r.finish(running[k]);
      running.erase(running.begin() + std::ptrdiff_t(k));
    }
  }
  for (size_t i = 0; i < order.size(); ++i) CHECK(order[i] == i + 1, "admission: requests start in arrival order");
  ok("admission");
}

// The host loops: every element of every loop is visited exactly once when loops run from several
// threads at once and from inside other loops' bodies, and a body's exception reaches the caller.
static void host_loops() {
  auto loop_sum = [](size_t n, size_t grain, bool nested) {
    std::vector<uint64_t> part((n + grain - 1) / grain, 0);
    parallel_chunks(n, grain, [&](size_t c, size_t b, size_t e) {
      uint64_t t = 0;
      for (size_t i = b; i < e; ++i) t += i;
      if (nested && c % 7 == 0) {
        std::vector<uint64_t> inner(8, 0);
        parallel_chunks(800, 100, [&](size_t ic, size_t ib, size_t ie) {
          for (size_t i = ib; i < ie; ++i) inner[ic] += 1;
        });
        for (uint64_t v : inner) t += v - 100;  // adds zero when every inner element ran once
      }
      part[c] = t;
    });
    uint64_t total = 0;
    for (uint64_t v : part) total += v;
    return total;
  };
  std::vector<std::thread> callers;
  std::atomic<int> bad{0};
  for (int t = 0; t < 4; ++t)
    callers.emplace_back([&, t] {
      for (int r = 0; r < 200; ++r) {
        const size_t n = 1000 + size_t(r * 37 + t * 11);
        if (loop_sum(n, 16 + size_t(r % 5), (r + t) % 3 == 0) != uint64_t(n) * (n - 1) / 2) bad.fetch_add(1);
      }
    });
  for (auto& c : callers) c.join();
  CHECK(bad.load() == 0, "concurrent and nested host loops");
  bool caught = false;
  try {
    parallel_chunks(1000, 10, [](size_t c, size_t, size_t) {
      if (c == 57) throw Error("test", "chunk 57");
    });
  } catch (const Error& e) {
    caught = e.message == "chunk 57";
  }
  CHECK(caught, "a body's exception reaches the caller");
  CHECK(loop_sum(5000, 50, false) == 5000ull * 4999 / 2, "the pool works after an exception");
  ok("host_loops");
}

int main() {
  host_loops();
  mt19937();
  regular_sampling();
  shrink();
  deriche();
  hdf5_container();
  itk_composite();
  convergence();
  wait_timeout();
  tuning();
  optimizer();
  speculative_search();
  jacobians();
  thresholds();
  taps();
  preprocess();
  mi_finish();
  admission();
  fastdiv();
  request_loop();
  std::printf("all host checks passed\n");
  return 0;
}
