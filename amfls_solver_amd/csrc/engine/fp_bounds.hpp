// Outward-rounded binary64 interval arithmetic for the certified validator.
//
// Same contract as the reference's floating-point bounds
// (src/algorithms/mathematics/floating_point_bounds.hpp:12-403): every
// operation is evaluated in x87 long double and stepped one binary64 ulp
// outward, gamma_k = k u / (1 - k u), and the norm interval is built from
// (max |v|, scaled square sum, length) -- the three numbers the device norm
// kernels return.
#pragma once

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <cstdint>
#include <limits>

namespace amflsx::fp {

struct Interval {
    double lower = 0.0;
    double upper = std::numeric_limits<double>::infinity();
};

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kDenormMin = std::numeric_limits<double>::denorm_min();
constexpr double kDblMax = std::numeric_limits<double>::max();

inline bool gradual_underflow_is_active() noexcept {
    volatile double tiny = kDenormMin;
    volatile double one = 1.0;
    volatile double kept = tiny * one;
    volatile double smallest = std::numeric_limits<double>::min();
    volatile double half = 0.5;
    volatile double produced = smallest * half;
    return kept == kDenormMin && produced > 0.0;
}

inline bool environment_ok() noexcept {
    return std::numeric_limits<double>::is_iec559 && std::fegetround() == FE_TONEAREST &&
           gradual_underflow_is_active();
}

inline double up_ld(long double v) noexcept {
    if (!(v >= 0.0L) || !std::isfinite(v) || v > static_cast<long double>(kDblMax)) return kInf;
    if (v == 0.0L) return 0.0;
    double r = static_cast<double>(v);
    if (static_cast<long double>(r) < v) r = std::nextafter(r, kInf);
    return std::nextafter(r, kInf);
}

inline double down_ld(long double v) noexcept {
    if (!(v > 0.0L) || !std::isfinite(v)) return 0.0;
    if (v > static_cast<long double>(kDblMax)) return kDblMax;
    double r = static_cast<double>(v);
    if (static_cast<long double>(r) > v) r = std::nextafter(r, 0.0);
    return std::nextafter(r, 0.0);
}

inline double gamma64(long long ops) noexcept {
    if (ops < 0) return kInf;
    const long double p = static_cast<long double>(ops) * (0.5L * std::numeric_limits<double>::epsilon());
    if (!(p < 1.0L)) return kInf;
    return up_ld(p / (1.0L - p));
}

inline double up_add(double a, double b) noexcept {
    return up_ld(static_cast<long double>(a) + static_cast<long double>(b));
}
inline double up_mul(double a, double b) noexcept {
    if (a < 0.0 || b < 0.0) return kInf;
    const long double p = static_cast<long double>(a) * static_cast<long double>(b);
    if (a > 0.0 && b > 0.0 && p == 0.0L) return kDenormMin;
    return up_ld(p);
}
inline double down_add(double a, double b) noexcept {
    if (a < 0.0 || b < 0.0 || !std::isfinite(a) || !std::isfinite(b)) return 0.0;
    return down_ld(static_cast<long double>(a) + static_cast<long double>(b));
}
inline double down_mul(double a, double b) noexcept {
    if (a < 0.0 || b < 0.0 || !std::isfinite(a) || !std::isfinite(b)) return 0.0;
    return down_ld(static_cast<long double>(a) * static_cast<long double>(b));
}
inline double down_div(double n, double d) noexcept {
    if (!(n > 0.0) || !(d > 0.0) || !std::isfinite(n) || !std::isfinite(d)) return 0.0;
    return down_ld(static_cast<long double>(n) / static_cast<long double>(d));
}
inline double down_diff(double a, double b) noexcept {
    if (!(a > b) || !std::isfinite(a) || b < 0.0 || !std::isfinite(b)) return 0.0;
    return down_ld(static_cast<long double>(a) - static_cast<long double>(b));
}
inline double up_div(double n, double d) noexcept {
    if (n == 0.0) return 0.0;
    if (!(d > 0.0) || n < 0.0 || !std::isfinite(n) || !std::isfinite(d)) return kInf;
    const long double q = static_cast<long double>(n) / static_cast<long double>(d);
    if (q == 0.0L) return kDenormMin;
    return up_ld(q);
}

inline double abs_roundoff(long long ops) noexcept {
    return up_ld(static_cast<long double>(ops) * static_cast<long double>(kDenormMin));
}

inline double pos_sum_upper(double stored, long long ops) noexcept {
    if (stored < 0.0 || !std::isfinite(stored) || ops < 0) return kInf;
    if (stored == 0.0) return 0.0;
    const double den = down_diff(1.0, gamma64(ops));
    if (!(den > 0.0)) return kInf;
    return up_div(up_add(stored, abs_roundoff(ops)), den);
}
inline double pos_fma_square_sum_upper(double stored, long long ops) noexcept {
    if (stored < 0.0 || !std::isfinite(stored) || ops < 0) return kInf;
    const double den = down_diff(1.0, gamma64(ops));
    if (!(den > 0.0)) return kInf;
    return up_div(up_add(stored, abs_roundoff(ops)), den);
}
inline double pos_sum_lower(double stored, long long ops) noexcept {
    if (!(stored > 0.0) || !std::isfinite(stored) || ops < 0) return 0.0;
    const double den = up_add(1.0, gamma64(ops));
    if (!std::isfinite(den)) return 0.0;
    return down_div(down_diff(stored, abs_roundoff(ops)), den);
}

inline double up_sqrt(double v) noexcept {
    if (v < 0.0 || !std::isfinite(v)) return v == kInf ? v : kInf;
    return up_ld(std::sqrt(static_cast<long double>(v)) *
                 (1.0L + 2.0L * std::numeric_limits<long double>::epsilon()));
}
inline double down_sqrt(double v) noexcept {
    if (!(v > 0.0) || !std::isfinite(v)) return 0.0;
    return down_ld(std::sqrt(static_cast<long double>(v)) *
                   (1.0L - 2.0L * std::numeric_limits<long double>::epsilon()));
}

inline double up_scale2(double v, int e) noexcept {
    if (v == 0.0) return 0.0;
    if (!(v > 0.0) || !std::isfinite(v)) return kInf;
    const double s = std::scalbn(v, e);
    if (std::isinf(s)) return s;
    if (s == 0.0) return kDenormMin;
    return std::nextafter(s, kInf);
}
inline double down_scale2(double v, int e) noexcept {
    if (!(v > 0.0) || !std::isfinite(v)) return 0.0;
    const double s = std::scalbn(v, e);
    if (std::isinf(s)) return kDblMax;
    if (s == 0.0) return 0.0;
    return std::nextafter(s, 0.0);
}

// Interval for ||v||_2 from the device statistics of v.
inline Interval norm_interval(double maxabs, double scaled_sum, double nonfinite,
                              long long count) noexcept {
    if (!environment_ok() || nonfinite > 0.0 || !std::isfinite(maxabs) ||
        !std::isfinite(scaled_sum))
        return {};
    if (maxabs == 0.0) return {0.0, 0.0};
    int e = 0;
    (void)std::frexp(maxabs, &e);
    const double lo2 = pos_sum_lower(scaled_sum, count);
    const double hi2 = pos_sum_upper(scaled_sum, count);
    const double slo = down_sqrt(lo2);
    const double shi = up_sqrt(hi2);
    const double serr = up_ld(std::sqrt(static_cast<long double>(count)) *
                              static_cast<long double>(kDenormMin));
    return {down_scale2(down_diff(slo, serr), e), up_scale2(up_add(shi, serr), e)};
}

// Nominal ||v||_2 from the same statistics (NaN when a value is nonfinite).
inline double nominal_norm(double maxabs, double scaled_sum, double nonfinite) noexcept {
    if (nonfinite > 0.0) return std::numeric_limits<double>::quiet_NaN();
    if (maxabs == 0.0) return 0.0;
    int e = 0;
    (void)std::frexp(maxabs, &e);
    return std::ldexp(std::sqrt(scaled_sum), e);
}

}  // namespace amflsx::fp
