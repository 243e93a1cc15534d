// Batched SyN on the device: several subjects registered to one template (the cohort form the
// reference runs as separate registrations, crates/fastwarp-cli/src/engine.rs:1016-1068 and
// python/fastwarp/_api.py:1104-1137). The subjects' SyN iterations run in lockstep and every step
// of an iteration is one launch over all of them (warps, CC pairs, channel weighting, update
// smoothing, update scaling, compositions, inversions); the template's smoothed levels are built
// once and read by every subject. Each subject's results are the bits of its own syn_fields run:
// the kernels give every item the arithmetic of its own launch, and each subject keeps its own
// decisions (convergence, inversion stops).
#pragma once

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

// register_resident for every subject against one template: the template is preprocessed, its
// centre of mass taken and its pyramid smoothed once; each subject's linear chain runs as in
// register_resident; the SyN stages run batched, kBatchItems subjects at a time.
std::vector<RegisterOutput> register_batch(HipBackend& b, HipLinearEngine& engine, const RegisterPlan& plan,
                                           const std::vector<const HImage*>& fixed,
                                           const std::vector<std::vector<const HImage*>>& moving,
                                           const volatile int32_t* cancel, std::vector<std::optional<HField>>* forward,
                                           std::vector<std::optional<HField>>* inverse);

}  // namespace fw::hip
