// Learned width controller: the 12 -> 64 -> 64 -> 1 MLP that can replace
// the reference's widening predicate (amfls.cpp:322-358, the cost gate of
// amfls_internal.hpp:318-324).  Weights live on the device; inference and
// training run the HIP kernels of mlp.hip.  The FP64 certificate remains the
// only termination test: the model only chooses whether to widen.
#pragma once

#include <array>
#include <vector>

#include "context.hpp"
#include "solver.hpp"

namespace amflsx {

// The 12 model inputs of a widening decision: log2 of the active width, the
// last four log10 certificate decreases, log10 of the last four per-level
// interval seconds (1e-9 floor), log10 of rows, columns and nonzeros.
std::array<float, kMlpInputs> width_features(const WidthFeatures& f);

class WidthModel {
public:
    // params: kMlpParams values; norm: 2 * kMlpInputs values (mean, inverse std).
    WidthModel(const float* params, const float* norm, Context& ctx);
    void set_params(const float* params);
    void params(float* out);
    void logits(const float* X_host, int64_t n, float* out_host);
    // Mean BCE and its gradient for a device batch (X_dev n x 12, y_dev n).
    float gradient(const float* X_dev, const float* y_dev, int64_t n, float* grad_host);
    // One Adam step on a device batch; returns the batch loss before the step.
    float train_step(const float* X_dev, const float* y_dev, int64_t n, float lr);
    // Per-feature box of the inputs the model was trained on.  A decision
    // whose inputs leave the box is taken by the heuristic (f.heuristic):
    // the model never extrapolates.  Default: unbounded.
    void set_domain(const float* lo, const float* hi);
    bool in_domain(const std::array<float, kMlpInputs>& x) const;
    bool widen(const WidthFeatures& f);
    // Decisions taken by the model / handed to the heuristic since creation.
    long long model_decisions() const { return model_decisions_; }
    long long fallback_decisions() const { return fallback_decisions_; }

private:
    Context& ctx_;
    DevBuf<float> params_, norm_, m_, v_, grad_, in_, out_;
    std::array<float, kMlpInputs> lo_, hi_;
    int step_ = 0;
    long long model_decisions_ = 0, fallback_decisions_ = 0;
};

}  // namespace amflsx
