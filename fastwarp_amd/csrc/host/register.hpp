// One full normalisation as register_stages_async runs it (crates/fastwarp-reg/src/stages.rs:286-403
// with register.rs:201-248): preprocess every channel (winsorise, window to [0, 1]), centre of mass
// on the raw first channel, Rigid then Affine Mattes MI on the first channel, then SyN on all
// channels from the collapsed affine. Generic over the SyN backend and the linear engine.
#pragma once

#include <chrono>

#include "linear.hpp"
#include "preprocess.hpp"
#include "syn.hpp"

namespace fw {

// stages.rs:74-86 `InitialTransform`: where the chain starts.
enum class InitialKind { CentreOfMass, Identity, Affine };

struct RegisterPlan {
  LinearRecipe rigid, affine;
  SynRecipe syn;
  float winsorize_lower = 0.005f, winsorize_upper = 0.995f;  // register.rs WINSORIZE_LOWER/UPPER
  bool run_linear = true, run_syn = true;
  InitialKind initial_kind = InitialKind::CentreOfMass;
  AffineTransform initial_affine;  // InitialKind::Affine

  RegisterPlan() {
    rigid.kind = LinearKind::Rigid;
    affine.kind = LinearKind::Affine;
  }
};

struct RegisterOutput {
  AffineTransform initial, total;
  LinearChainResult linear;
  SynFields syn;
  double seconds_preprocess = 0.0, seconds_linear = 0.0, seconds_syn = 0.0, seconds_total = 0.0;
};

template <typename B, typename E, typename V>
RegisterOutput register_full(B& backend, E& engine, const RegisterPlan& plan, const std::vector<V>& fixed,
                             const std::vector<V>& moving, const volatile int32_t* cancel, FieldOutputs outputs = {}) {
  using clock = std::chrono::steady_clock;
  auto secs = [](clock::time_point a) { return std::chrono::duration<double>(clock::now() - a).count(); };
  const auto t_all = clock::now();
  if (fixed.empty() || fixed.size() != moving.size())
    invalid("register", "fixed and moving must have the same, non-zero number of channels");
  if (plan.run_syn && plan.syn.channel_weights.size() != fixed.size())
    invalid("stage plan", "the number of channels differs from the number of channel weights");
  {
    HostTimer t_check("validate");
    bool finite = true;
    for (const auto* set : {&fixed, &moving})
      for (const V& v : *set) {
        std::atomic<bool> ok{true};
        parallel_for(v.data.size(), [&](size_t b, size_t e) {
          for (size_t i = b; i < e; ++i)
            if (!std::isfinite(v.data[i])) {
              ok = false;
              return;
            }
        });
        finite = finite && ok;
      }
    if (!finite) invalid("register", "an input image has non-finite (NaN or infinite) voxels");
  }
  RegisterOutput out;
  auto t = clock::now();
  std::vector<SynChannel> channels;
  for (size_t c = 0; c < fixed.size(); ++c)
    channels.push_back({preprocess_volume(fixed[c], plan.winsorize_lower, plan.winsorize_upper),
                        preprocess_volume(moving[c], plan.winsorize_lower, plan.winsorize_upper)});
  out.seconds_preprocess = secs(t);
  std::vector<std::pair<const Image*, const Image*>> pre;
  for (const SynChannel& c : channels) pre.push_back({&c.fixed, &c.moving});
  SmoothCache<B> cache(backend, std::move(pre));
  t = clock::now();
  const AffineTransform start = plan.initial_kind == InitialKind::CentreOfMass ? initial_transform_com(fixed[0], moving[0])
                                : plan.initial_kind == InitialKind::Identity   ? AffineTransform{}
                                                                                : plan.initial_affine;
  if (plan.run_linear) {
    out.linear = linear_chain_of(engine, plan.rigid, plan.affine, start, cache, cancel);
    out.initial = out.linear.initial;
    out.total = out.linear.total;
  } else {
    out.initial = out.total = start;
  }
  out.seconds_linear = secs(t);
  if (plan.run_syn) {
    backend.release_idle_buffers();
    t = clock::now();
    out.syn = syn_stage_of(backend, plan.syn, out.total, cache, cancel, outputs, nullptr, nullptr, SynStageInputs<B>{});
    out.seconds_syn = secs(t);
  }
  out.seconds_total = secs(t_all);
  return out;
}

// The same normalisation from raw images already on the device: every channel is preprocessed
// and the centre of mass taken on the device, and the fields stay there (forward_device,
// inverse_device, when given). The backend provides preprocess_image and centre_of_gravity.
template <typename B, typename E>
RegisterOutput register_resident(B& backend, E& engine, const RegisterPlan& plan,
                                 const std::vector<const typename B::ImageH*>& fixed,
                                 const std::vector<const typename B::ImageH*>& moving, const volatile int32_t* cancel,
                                 std::optional<typename B::FieldH>* forward_device = nullptr,
                                 std::optional<typename B::FieldH>* inverse_device = nullptr) {
  using clock = std::chrono::steady_clock;
  using ImageH = typename B::ImageH;
  auto secs = [](clock::time_point a) { return std::chrono::duration<double>(clock::now() - a).count(); };
  const auto t_all = clock::now();
  if (fixed.empty() || fixed.size() != moving.size())
    invalid("register", "fixed and moving must have the same, non-zero number of channels");
  if (plan.run_syn && plan.syn.channel_weights.size() != fixed.size())
    invalid("stage plan", "the number of channels differs from the number of channel weights");
  for (size_t c = 0; c < fixed.size(); ++c)
    if (!fixed[c] || !moving[c]) invalid("register", "missing image");
  RegisterOutput out;
  auto t = clock::now();
  std::vector<typename SmoothCache<B>::Pair> resident;
  for (size_t c = 0; c < fixed.size(); ++c)
    resident.push_back({std::make_shared<const ImageH>(backend.preprocess_image(*fixed[c], plan.winsorize_lower, plan.winsorize_upper)),
                        std::make_shared<const ImageH>(backend.preprocess_image(*moving[c], plan.winsorize_lower, plan.winsorize_upper))});
  out.seconds_preprocess = secs(t);
  SmoothCache<B> cache(backend, std::move(resident));
  t = clock::now();
  AffineTransform start = plan.initial_affine;
  if (plan.initial_kind == InitialKind::CentreOfMass) {
    const Vec3 fc = backend.centre_of_gravity(*fixed[0]), mc = backend.centre_of_gravity(*moving[0]);
    start = AffineTransform{kIdentity3, vsub(mc, fc), fc};
  } else if (plan.initial_kind == InitialKind::Identity) {
    start = AffineTransform{};
  }
  if (plan.run_linear) {
    out.linear = linear_chain_of(engine, plan.rigid, plan.affine, start, cache, cancel);
    out.initial = out.linear.initial;
    out.total = out.linear.total;
  } else {
    out.initial = out.total = start;
  }
  out.seconds_linear = secs(t);
  if (plan.run_syn) {
    backend.release_idle_buffers();
    t = clock::now();
    out.syn = syn_stage_of(backend, plan.syn, out.total, cache, cancel, FieldOutputs{}, forward_device, inverse_device,
                           SynStageInputs<B>{});
    out.seconds_syn = secs(t);
  }
  out.seconds_total = secs(t_all);
  return out;
}

}  // namespace fw
