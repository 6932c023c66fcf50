#include "width_model.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

namespace amflsx {

std::array<float, kMlpInputs> width_features(const WidthFeatures& f) {
    std::array<float, kMlpInputs> x{};
    x[0] = static_cast<float>(std::log2(std::max(1, f.active_width)));
    for (int i = 0; i < 4; ++i) x[1 + i] = static_cast<float>(std::isfinite(f.bound_decrease[i]) ? f.bound_decrease[i] : 0.0);
    for (int i = 0; i < 4; ++i) x[5 + i] = static_cast<float>(std::log10(std::max(f.step_seconds[i], 1e-9)));
    x[9] = static_cast<float>(std::log10(std::max(f.rows, 1.0)));
    x[10] = static_cast<float>(std::log10(std::max(f.cols, 1.0)));
    x[11] = static_cast<float>(std::log10(std::max(f.nnz, 1.0)));
    return x;
}

WidthModel::WidthModel(const float* params, const float* norm, Context& ctx) : ctx_(ctx) {
    params_.reset(kMlpParams);
    norm_.reset(2 * kMlpInputs);
    m_.reset(kMlpParams);
    v_.reset(kMlpParams);
    grad_.reset(kMlpParams + 1);
    in_.reset(kMlpInputs);
    out_.reset(1);
    lo_.fill(-std::numeric_limits<float>::infinity());
    hi_.fill(std::numeric_limits<float>::infinity());
    AMFLS_HIP_CHECK(hipMemcpy(params_.get(), params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    AMFLS_HIP_CHECK(hipMemcpy(norm_.get(), norm, 2 * kMlpInputs * sizeof(float), hipMemcpyHostToDevice));
    AMFLS_HIP_CHECK(hipMemset(m_.get(), 0, kMlpParams * sizeof(float)));
    AMFLS_HIP_CHECK(hipMemset(v_.get(), 0, kMlpParams * sizeof(float)));
}

void WidthModel::set_params(const float* params) {
    AMFLS_HIP_CHECK(hipMemcpy(params_.get(), params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
}

void WidthModel::params(float* out) {
    ctx_.download(out, params_.get(), kMlpParams);
}

void WidthModel::logits(const float* X_host, int64_t n, float* out_host) {
    if (n <= 0) return;
    // buffers grow once and are reused, so a decision inside a solve does
    // not allocate (an allocation would also wait for the device)
    in_.ensure(static_cast<size_t>(n) * kMlpInputs);
    out_.ensure(static_cast<size_t>(n));
    AMFLS_HIP_CHECK(hipMemcpyAsync(in_.get(), X_host, static_cast<size_t>(n) * kMlpInputs * sizeof(float),
                                   hipMemcpyHostToDevice, ctx_.stream));
    mlp_forward(ctx_.stream, params_.get(), norm_.get(), n, in_.get(), out_.get());
    ctx_.download(out_host, out_.get(), static_cast<size_t>(n));
}

void WidthModel::set_domain(const float* lo, const float* hi) {
    for (int i = 0; i < kMlpInputs; ++i) {
        if (std::isnan(lo[i]) || std::isnan(hi[i]) || lo[i] > hi[i])
            throw std::invalid_argument("width model domain needs lo <= hi for every input");
        lo_[i] = lo[i];
        hi_[i] = hi[i];
    }
}

bool WidthModel::in_domain(const std::array<float, kMlpInputs>& x) const {
    for (int i = 0; i < kMlpInputs; ++i)
        if (!(x[i] >= lo_[i] && x[i] <= hi_[i])) return false;
    return true;
}

float WidthModel::gradient(const float* X_dev, const float* y_dev, int64_t n, float* grad_host) {
    mlp_gradient(ctx_.stream, params_.get(), norm_.get(), n, X_dev, y_dev, grad_.get(), grad_.get() + kMlpParams,
                  ctx_.scratch);
    std::vector<float> g(kMlpParams + 1);
    ctx_.download(g.data(), grad_.get(), g.size());
    if (grad_host) std::copy(g.begin(), g.begin() + kMlpParams, grad_host);
    return g[kMlpParams];
}

float WidthModel::train_step(const float* X_dev, const float* y_dev, int64_t n, float lr) {
    mlp_gradient(ctx_.stream, params_.get(), norm_.get(), n, X_dev, y_dev, grad_.get(), grad_.get() + kMlpParams,
                 ctx_.scratch);
    mlp_adam(ctx_.stream, params_.get(), m_.get(), v_.get(), grad_.get(), lr, ++step_);
    float loss = 0.0f;
    ctx_.download(&loss, grad_.get() + kMlpParams, 1);
    return loss;
}

bool WidthModel::widen(const WidthFeatures& f) {
    const auto x = width_features(f);
    if (!in_domain(x)) {
        ++fallback_decisions_;
        return f.heuristic;
    }
    ++model_decisions_;
    float z = 0.0f;
    logits(x.data(), 1, &z);
    return z > 0.0f;
}

}  // namespace amflsx
