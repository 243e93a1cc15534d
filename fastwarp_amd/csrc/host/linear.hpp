// The linear stages (crates/fastwarp-reg/src/linear.rs, init.rs): per level the fixed grid shrunk
// by the level factors is the virtual domain, both full-resolution images are smoothed, one
// REGULAR sample set is drawn with the next seed of the stage, Mattes MI is initialised on it and
// ITK's conjugate gradient runs. Generic over a LinearEngine whose prepare_level(moving, points,
// fixed_bin, binning) returns a LinearMetric.
#pragma once

#include <future>
#include <map>
#include <memory>
#include <optional>

#include "interp.hpp"
#include "mi_host.hpp"
#include "optimizer.hpp"
#include "pyramid.hpp"
#include "sampling.hpp"

namespace fw {

enum class LinearKind { Rigid, Affine };

// linear.rs:75-145 `LinearRecipe`
struct LinearRecipe {
  LinearKind kind = LinearKind::Rigid;
  double gradient_step = 0.1;
  std::vector<uint32_t> iterations{1000, 500, 250, 100};
  double convergence_threshold = 1e-6;
  size_t convergence_window = 10;
  std::vector<uint32_t> shrink_factors{8, 4, 2, 1};
  std::vector<double> smoothing_sigmas{3.0, 2.0, 1.0, 0.0};
  SigmaUnits sigma_units = SigmaUnits::Voxels;
  size_t bins = 32;
  float sampling_percentage = 0.25f;
  uint32_t seed = 20260910u;

  size_t levels() const { return iterations.size(); }
  // linear.rs:191-239 `validate`
  void validate() const {
    const size_t n = iterations.size();
    if (n == 0 || shrink_factors.size() != n || smoothing_sigmas.size() != n)
      invalid("linear recipe", "iterations, shrink factors and smoothing sigmas must have one entry per level");
    for (auto f : shrink_factors)
      if (f == 0) invalid("linear recipe", "iterations, shrink factors and smoothing sigmas must have one entry per level");
    if (convergence_window < 2 || bins <= 4) invalid("linear recipe", "window below 2 or at most 4 bins");
    if (!(sampling_percentage > 0.0f && sampling_percentage <= 1.0f)) invalid("linear recipe", "sampling percentage outside (0, 1]");
    if (!(std::isfinite(gradient_step) && gradient_step > 0.0)) invalid("linear recipe", "gradient step must be positive");
    for (double s : smoothing_sigmas)
      if (!(std::isfinite(s) && s >= 0.0)) invalid("linear recipe", "smoothing sigmas must be finite and non-negative");
  }
};

// Smoothed full-resolution image pairs shared by the rigid, affine and SyN stages of one
// registration (the reference's SmoothCache, linear.rs:569-754), keyed by channel and sigma and
// held as backend images (device-resident on the HIP backend). Each channel's preprocessed pair is
// uploaded once, on first use; sigma <= 0 is the upload itself.
template <typename B>
class SmoothCache {
 public:
  using ImageH = typename B::ImageH;
  using Ptr = std::shared_ptr<const ImageH>;
  struct Pair {
    Ptr fixed, moving;
  };

  // Host images, uploaded on first use.
  SmoothCache(B& backend, std::vector<std::pair<const Image*, const Image*>> channels)
      : b_(backend), host_(std::move(channels)) {
    for (const auto& c : host_) grids_.push_back(c.first->grid);
  }
  // Images already on the device (the unsmoothed level of every channel).
  SmoothCache(B& backend, std::vector<Pair> resident) : b_(backend) {
    for (size_t c = 0; c < resident.size(); ++c) {
      grids_.push_back(resident[c].fixed->grid);
      pairs_[{c, 0.0}] = resident[c];
    }
  }

  B& backend() { return b_; }
  const Grid& fixed_grid(size_t channel) const { return grids_.at(channel); }

  // Smoothed fixed images shared between the caches of one template's subjects (each cache holds
  // the same unsmoothed fixed images): a level of the template is smoothed once, by whichever
  // cache asks for it first, and the others take that image. Caches used from several threads
  // need the map complete beforehand (they only read it then).
  using SharedFixed = std::map<std::pair<size_t, double>, Ptr>;
  void share_fixed(std::shared_ptr<SharedFixed> shared) { shared_ = std::move(shared); }
  size_t channels() const { return grids_.size(); }

  Pair get(size_t channel, double sigma, SigmaUnits units) {
    if (channel >= grids_.size()) invalid("smoothing", "no such channel");
    const double key = sigma > 0.0 ? sigma : 0.0;
    auto it = pairs_.find({channel, key});
    if (it != pairs_.end()) return it->second;
    Pair p;
    if (key == 0.0) {
      p.fixed = std::make_shared<const ImageH>(b_.upload_image(*host_[channel].first));
      p.moving = std::make_shared<const ImageH>(b_.upload_image(*host_[channel].second));
    } else {
      const Pair raw = get(channel, 0.0, units);
      HostTimer t("smooth_images");
      if (shared_) {
        // several threads may read a complete map; an entry is only added by a single user
        const auto hit = shared_->find({channel, key});
        if (hit != shared_->end() && hit->second) {
          p.fixed = hit->second;
        } else {
          p.fixed = std::make_shared<const ImageH>(b_.smooth_image(*raw.fixed, sigma, units));
          (*shared_)[{channel, key}] = p.fixed;
        }
      } else {
        p.fixed = std::make_shared<const ImageH>(b_.smooth_image(*raw.fixed, sigma, units));
      }
      p.moving = std::make_shared<const ImageH>(b_.smooth_image(*raw.moving, sigma, units));
    }
    pairs_[{channel, key}] = p;
    return p;
  }

  // Finite range of a cached moving image (mi.rs:209-236 reads it for the moving bins).
  std::pair<double, double> moving_range(size_t channel, double sigma, SigmaUnits units) {
    const double key = sigma > 0.0 ? sigma : 0.0;
    auto it = ranges_.find({channel, key});
    if (it != ranges_.end()) return it->second;
    const Pair p = get(channel, sigma, units);
    double lo = 0.0, hi = 0.0;
    if (!b_.finite_range(*p.moving, lo, hi)) invalid("mattes", "the moving image has no finite voxel");
    ranges_[{channel, key}] = {lo, hi};
    return {lo, hi};
  }

 private:
  B& b_;
  std::vector<std::pair<const Image*, const Image*>> host_;
  std::vector<Grid> grids_;
  std::map<std::pair<size_t, double>, Pair> pairs_;
  std::map<std::pair<size_t, double>, std::pair<double, double>> ranges_;
  std::shared_ptr<SharedFixed> shared_;
};

// mi.rs:193-207 `fixed_sample_range`: count and range of the finite nearest-voxel values.
inline size_t fixed_sample_range(const HostVec<double>& nearest, double& lo, double& hi) {
  constexpr size_t kGrain = size_t(1) << 14;
  const size_t chunks = (nearest.size() + kGrain - 1) / kGrain;
  struct Part {
    size_t count = 0;
    double lo = 0, hi = 0;
  };
  std::vector<Part> parts(chunks);
  parallel_chunks(nearest.size(), kGrain, [&](size_t c, size_t b, size_t e) {
    Part p;
    for (size_t i = b; i < e; ++i) {
      const double v = nearest[i];
      if (!std::isfinite(v)) continue;
      if (p.count == 0) p.lo = p.hi = v;
      else {
        p.lo = std::min(p.lo, v);
        p.hi = std::max(p.hi, v);
      }
      ++p.count;
    }
    parts[c] = p;
  });
  size_t count = 0;
  for (const Part& p : parts) {
    if (p.count == 0) continue;
    if (count == 0) {
      lo = p.lo;
      hi = p.hi;
    } else {
      lo = std::min(lo, p.lo);
      hi = std::max(hi, p.hi);
    }
    count += p.count;
  }
  return count;
}

struct PreparedSamples {
  HostVec<Vec3> points;
  HostVec<uint32_t> fixed_bin;
  MiBinning binning;
};

// linear.rs:251-279 `prepare_samples` with MiBinning::new (mi.rs:209-236), from the fixed image's
// nearest-voxel and linear values at the raw sample points (NaN outside) and the moving range:
// samples whose linear fixed value is not finite are dropped, the rest get their fixed bin.
inline PreparedSamples prepare_samples(const HostVec<Vec3>& samples, const HostVec<double>& nearest,
                                       const HostVec<double>& linear, std::pair<double, double> moving_range,
                                       size_t bins) {
  HostTimer host_timer("prepare_samples");
  if (bins <= 4) invalid("mattes", "at least five histogram bins are required");
  double flo = 0, fhi = 0;
  if (fixed_sample_range(nearest, flo, fhi) == 0) invalid("mattes", "no fixed-image samples were observed");
  if (fhi - flo <= kEps) invalid("mattes", "the fixed image is constant over the samples");
  PreparedSamples out;
  out.binning = MiBinning::from_ranges(bins, flo, fhi, moving_range.first, moving_range.second);
  // keep the samples whose linear value is finite, in order: count per chunk, then fill each
  // chunk's slice of the output in parallel
  constexpr size_t kGrain = size_t(1) << 14;
  const size_t m = samples.size();
  const size_t chunks = (m + kGrain - 1) / kGrain;
  std::vector<size_t> kept(chunks + 1, 0);
  parallel_chunks(m, kGrain, [&](size_t c, size_t b, size_t e) {
    size_t k = 0;
    for (size_t i = b; i < e; ++i) k += std::isfinite(linear[i]) ? 1 : 0;
    kept[c + 1] = k;
  });
  for (size_t c = 0; c < chunks; ++c) kept[c + 1] += kept[c];
  out.points.resize(kept[chunks]);
  out.fixed_bin.resize(kept[chunks]);
  parallel_chunks(m, kGrain, [&](size_t c, size_t b, size_t e) {
    size_t o = kept[c];
    for (size_t i = b; i < e; ++i)
      if (std::isfinite(linear[i])) {
        out.points[o] = samples[i];
        out.fixed_bin[o] = uint32_t(out.binning.fixed_bin(linear[i]));
        ++o;
      }
  });
  return out;
}

struct LinearLevelTrace {
  size_t level = 0;
  std::array<uint32_t, 3> shrink_factors{};
  std::array<size_t, 3> grid_size{};
  size_t samples = 0;
  CgTrace cg;
  double seconds = 0.0;
  std::vector<double> parameters;
  std::string skipped;
};

struct LinearStageResult {
  ParametricTransform transform;
  AffineTransform affine;
  std::vector<LinearLevelTrace> levels;
};

// The engine levels of a stage, kept so that a chained stage with the same sampling reuses them:
// the rigid and the affine stage of a chain draw identical sample sets per level (the same seed
// sequence over the same smoothed images), as the reference's sample cache exploits.
template <typename Engine>
struct LevelCache {
  std::map<size_t, std::shared_ptr<typename Engine::Level>> levels;
  std::map<size_t, size_t> sample_counts;
  std::map<size_t, std::string> skipped;
};

// linear.rs:873-1068 `run_linear_stage_cached_async`.
template <typename Engine, typename B>
LinearStageResult run_linear_stage(Engine& engine, const LinearRecipe& recipe, SmoothCache<B>& cache,
                                   const AffineTransform& outer, const ParametricTransform& start,
                                   LevelCache<Engine>* levels_cache, const volatile int32_t* cancel) {
  const Grid& fixed_grid = cache.fixed_grid(0);
  recipe.validate();
  if (start.count() != (recipe.kind == LinearKind::Rigid ? 6u : 12u))
    invalid("linear stage", "start transform does not match the stage kind");
  SeedSequence seeds(recipe.seed);
  LinearMovingTransform current{outer, start};
  LinearStageResult result;
  // Level l draws its sample points with seed + l whether or not it samples (each level takes one
  // seed, sampling.rs:159-163), so the later levels' points are drawn on worker threads while the
  // earlier levels optimise (the reference prefetches them on a thread too, linear.rs:773-869).
  auto level_grid = [&](size_t level) {
    return level_virtual_grid(fixed_grid, shrink_factors_per_dimension(recipe.shrink_factors[level], fixed_grid.spacing));
  };
  auto needs_samples = [&](size_t level) {
    return !(levels_cache && (levels_cache->skipped.count(level) || levels_cache->levels.count(level)));
  };
  auto draw = [&recipe](Grid grid, uint32_t seed) {
    SeedSequence s(seed);
    return metric_sample_points(grid, recipe.sampling_percentage, s);
  };
  // every level that samples starts drawing at once, each on its own worker thread
  std::vector<std::future<HostVec<Vec3>>> prefetched(recipe.levels());
  for (size_t level = 1; level < recipe.levels(); ++level)
    if (needs_samples(level))
      prefetched[level] = std::async(std::launch::async, draw, level_grid(level), uint32_t(recipe.seed + level));
  for (size_t level = 0; level < recipe.levels(); ++level) {
    check_cancel(cancel);
    const auto t0 = std::chrono::steady_clock::now();
    const auto factors = shrink_factors_per_dimension(recipe.shrink_factors[level], fixed_grid.spacing);
    const Grid grid = level_grid(level);
    std::optional<HostVec<Vec3>> ready;
    if (prefetched[level].valid()) {
      HostTimer wait("sample_wait");
      ready = prefetched[level].get();
    }
    const double sigma = recipe.smoothing_sigmas[level];
    LinearLevelTrace tr;
    tr.level = level;
    tr.shrink_factors = factors;
    tr.grid_size = grid.size;
    std::shared_ptr<typename Engine::Level> metric;
    if (levels_cache && levels_cache->skipped.count(level)) {
      seeds.next_seed();
      tr.skipped = levels_cache->skipped[level];
    } else if (levels_cache && levels_cache->levels.count(level)) {
      seeds.next_seed();
      metric = levels_cache->levels[level];
      tr.samples = levels_cache->sample_counts[level];
    } else {
      const auto images = cache.get(0, sigma, recipe.sigma_units);
      HostVec<Vec3> raw;
      if (ready) {
        seeds.next_seed();
        raw = std::move(*ready);
      } else {
        raw = metric_sample_points(grid, recipe.sampling_percentage, seeds);
      }
      HostVec<double> nearest, linear;
      cache.backend().sample_values(*images.fixed, raw, nearest, linear);
      double lo = 0, hi = 0;
      const size_t observed = fixed_sample_range(nearest, lo, hi);
      if (observed == 0) tr.skipped = "no sample point lands on the fixed image";
      else if (hi - lo <= kEps) tr.skipped = "the fixed image takes a single value at its sample points";
      if (tr.skipped.empty()) {
        PreparedSamples prep = prepare_samples(raw, nearest, linear, cache.moving_range(0, sigma, recipe.sigma_units), recipe.bins);
        if (prep.points.empty()) tr.skipped = "no sample point lies inside the fixed image";
        else {
          tr.samples = prep.points.size();
          metric = std::make_shared<typename Engine::Level>(
              engine.prepare_level(images.moving, prep.points, prep.fixed_bin, prep.binning));
        }
      }
      if (levels_cache) {
        if (metric) {
          levels_cache->levels[level] = metric;
          levels_cache->sample_counts[level] = tr.samples;
        } else {
          levels_cache->skipped[level] = tr.skipped;
        }
      }
    }
    if (metric) {
      CgSettings st;
      st.max_step_mm = recipe.gradient_step;
      st.iterations = recipe.iterations[level];
      st.convergence_threshold = recipe.convergence_threshold;
      st.convergence_window = recipe.convergence_window;
      auto [next, cg] = optimise_level(*metric, corner_points(grid), current, st, cancel);
      current = next;
      tr.cg = std::move(cg);
    }
    tr.parameters = current.optimised.parameters();
    tr.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    result.levels.push_back(std::move(tr));
  }
  bool all_skipped = !result.levels.empty();
  for (auto& l : result.levels) all_skipped = all_skipped && !l.skipped.empty();
  if (all_skipped)
    invalid("linear stage", "every level of the stage has too few usable samples for Mattes mutual information");
  result.transform = current.optimised;
  result.affine = current.optimised.to_affine();
  return result;
}

// init.rs:47-86: intensity-weighted centre (ITK ImageMomentsCalculator) and the initial
// translation moving centre - fixed centre about the fixed centre.
// Summed per plane on the host threads, the planes' sums then added in plane order.
template <typename V>
Vec3 centre_of_gravity(const V& vol) {
  HostTimer host_timer("centre_of_mass");
  const Grid& g = vol.grid;
  const size_t nx = g.size[0], ny = g.size[1], nz = g.size[2];
  struct Part {
    double mass = 0.0;
    Vec3 cg{0, 0, 0};
  };
  std::vector<Part> parts(nz);
  parallel_chunks(nz, 1, [&](size_t k, size_t, size_t) {
    Part p;
    for (size_t j = 0; j < ny; ++j) {
      const double* row = vol.data.data() + nx * (j + ny * k);
      for (size_t i = 0; i < nx; ++i) {
        const double v = row[i];
        if (v == 0.0) continue;
        const Vec3 x = g.index_to_point({double(i), double(j), double(k)});
        p.mass += v;
        for (int a = 0; a < 3; ++a) p.cg[a] += x[a] * v;
      }
    }
    parts[k] = p;
  });
  double mass = 0.0;
  Vec3 cg{0, 0, 0};
  for (const Part& p : parts) {
    mass += p.mass;
    for (int a = 0; a < 3; ++a) cg[a] += p.cg[a];
  }
  if (mass == 0.0) invalid("initial transform", "total mass of the image is zero");
  return {cg[0] / mass, cg[1] / mass, cg[2] / mass};
}

template <typename V>
AffineTransform initial_transform_com(const V& fixed, const V& moving) {
  const Vec3 fc = centre_of_gravity(fixed), mc = centre_of_gravity(moving);
  return AffineTransform{kIdentity3, vsub(mc, fc), fc};
}

struct LinearChainResult {
  AffineTransform initial;
  LinearStageResult rigid, affine;
  AffineTransform total;
};

// linear.rs:1148-1212 `run_linear_chain_cached_async`: centre of mass, Rigid, Affine, collapsed
// about the last linear transform's centre.
template <typename Engine, typename B>
LinearChainResult run_linear_chain(Engine& engine, const LinearRecipe& rigid, const LinearRecipe& affine,
                                   const AffineTransform& initial, SmoothCache<B>& cache, const volatile int32_t* cancel) {
  if (rigid.kind != LinearKind::Rigid || affine.kind != LinearKind::Affine)
    invalid("linear chain", "expected a rigid then an affine recipe");
  LinearChainResult out;
  out.initial = initial;
  const bool same_sampling = rigid.seed == affine.seed && rigid.shrink_factors == affine.shrink_factors &&
                             rigid.smoothing_sigmas == affine.smoothing_sigmas && rigid.bins == affine.bins &&
                             rigid.sampling_percentage == affine.sampling_percentage && rigid.sigma_units == affine.sigma_units;
  LevelCache<Engine> levels;
  out.rigid = run_linear_stage(engine, rigid, cache, out.initial, ParametricTransform::euler3d({0, 0, 0, 0, 0, 0}, out.initial.center),
                               same_sampling ? &levels : nullptr, cancel);
  const AffineTransform after_rigid = AffineTransform::compose(out.initial, out.rigid.affine);
  const ParametricTransform start = ParametricTransform::affine(AffineTransform{kIdentity3, {0, 0, 0}, out.rigid.affine.center});
  out.affine = run_linear_stage(engine, affine, cache, after_rigid, start, same_sampling ? &levels : nullptr, cancel);
  out.total = with_center(AffineTransform::compose(after_rigid, out.affine.affine), out.affine.affine.center);
  return out;
}

// The linear chain a registration runs with engine E: the chain above. An engine may provide its
// own chain for its type (an overload found by argument-dependent lookup).
template <typename Engine, typename B>
LinearChainResult linear_chain_of(Engine& engine, const LinearRecipe& rigid, const LinearRecipe& affine,
                                  const AffineTransform& initial, SmoothCache<B>& cache, const volatile int32_t* cancel) {
  return run_linear_chain(engine, rigid, affine, initial, cache, cancel);
}

// The same from the raw first channel: the centre-of-mass initialisation first.
template <typename Engine, typename B, typename V>
LinearChainResult run_linear_chain(Engine& engine, const LinearRecipe& rigid, const LinearRecipe& affine, const V& fixed_raw,
                                   const V& moving_raw, SmoothCache<B>& cache, const volatile int32_t* cancel) {
  return run_linear_chain(engine, rigid, affine, initial_transform_com(fixed_raw, moving_raw), cache, cancel);
}

}  // namespace fw
