// Batched SyN on the device: several subjects registered to one template (the cohort form the
// reference runs as separate registrations, crates/fastwarp-cli/src/engine.rs:1016-1068 and
// python/fastwarp/_api.py:1104-1137). The subjects' SyN iterations run in lockstep and every step
// of an iteration is one launch over all of them (warps, CC pairs, channel weighting, update
// smoothing, update scaling, compositions, inversions); the template's smoothed levels are built
// once and read by every subject. Each subject's results are the bits of its own syn_fields run:
// the kernels give every item the arithmetic of its own launch, and each subject keeps its own
// decisions (convergence, inversion stops).
#pragma once

#include <memory>
#include <optional>
#include <vector>

#include "../host/register.hpp"
#include "backend.hpp"

namespace fw::hip {

// syn_fields (crates/fastwarp-reg/src/syn.rs:776-986) of up to kBatchItems subjects, CC channels
// only. caches[s] holds subject s's channels (the fixed images must be one template, on one grid
// for every subject); initial[s] is its linear start. Fields stay on the device when forward /
// inverse are given (one entry per subject). An error of one subject fails the batch.
std::vector<SynFields> syn_fields_batch(HipBackend& b, const SynRecipe& recipe, const std::vector<AffineTransform>& initial,
                                        const std::vector<SmoothCache<HipBackend>*>& caches, const volatile int32_t* cancel,
                                        std::vector<std::optional<HField>>* forward,
                                        std::vector<std::optional<HField>>* inverse);

// Worker contexts of one device (each its own HIP stream and buffer pool), kept between calls, for
// linear fits that run side by side.
class WorkerPool {
 public:
  explicit WorkerPool(int device) : device_(device) {}
  // The first `count` workers, created as needed.
  std::vector<HipBackend*> take(size_t count);

 private:
  int device_;
  std::vector<std::unique_ptr<HipBackend>> workers_;
};

// One linear chain (centre of mass, Rigid, Affine on the first channel) of a fit: the raw images
// (device, finished) and, where a template is shared, its preprocessed image, its complete map of
// smoothed levels and its centre of mass.
struct LinearFitJob {
  const HImage* fixed_raw = nullptr;
  const HImage* moving_raw = nullptr;
  std::shared_ptr<const HImage> fixed_pre, moving_pre;
  std::shared_ptr<SmoothCache<HipBackend>::SharedFixed> shared;
  std::optional<Vec3> fixed_centre;
  LinearChainResult out;
  double seconds_preprocess = 0.0, seconds_linear = 0.0;
};

// The linear chains of `jobs` as register_resident runs them, spread over the workers: a worker
// takes the next job when it finishes one, so the fits' kernels and host round trips overlap on the
// device. Each job's result is the bits of its own run. The first error is raised once every
// worker has stopped.
void run_linear_fits(const std::vector<HipBackend*>& workers, const RegisterPlan& plan, std::vector<LinearFitJob>& jobs,
                     const volatile int32_t* cancel);

// register_resident for every subject against one template: the template is preprocessed, its
// centre of mass taken and its pyramid smoothed once; the subjects' linear chains run side by side
// on `workers`; the SyN stages run batched, kBatchItems subjects at a time.
std::vector<RegisterOutput> register_batch(HipBackend& b, const std::vector<HipBackend*>& workers, const RegisterPlan& plan,
                                           const std::vector<const HImage*>& fixed,
                                           const std::vector<std::vector<const HImage*>>& moving,
                                           const volatile int32_t* cancel, std::vector<std::optional<HField>>* forward,
                                           std::vector<std::optional<HField>>* inverse);

}  // namespace fw::hip
