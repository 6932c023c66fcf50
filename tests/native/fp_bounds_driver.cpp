// Test driver for the engine's outward-rounded bound arithmetic: reads
// "op a b" lines (hex floats) from stdin and prints each result as a hex
// float, so a test can compare it with exact rational arithmetic.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../amfls_solver_amd/csrc/engine/fp_bounds.hpp"

using namespace amflsx::fp;

int main() {
    char op[32];
    char sa[64], sb[64];
    if (!environment_ok()) {
        std::printf("environment\n");
        return 1;
    }
    while (std::scanf("%31s %63s %63s", op, sa, sb) == 3) {
        const double a = std::strtod(sa, nullptr), b = std::strtod(sb, nullptr);
        double r = 0.0, r2 = 0.0;
        bool pair = false;
        if (!std::strcmp(op, "up_add")) r = up_add(a, b);
        else if (!std::strcmp(op, "down_add")) r = down_add(a, b);
        else if (!std::strcmp(op, "up_mul")) r = up_mul(a, b);
        else if (!std::strcmp(op, "down_mul")) r = down_mul(a, b);
        else if (!std::strcmp(op, "up_div")) r = up_div(a, b);
        else if (!std::strcmp(op, "down_div")) r = down_div(a, b);
        else if (!std::strcmp(op, "down_diff")) r = down_diff(a, b);
        else if (!std::strcmp(op, "up_sqrt")) r = up_sqrt(a);
        else if (!std::strcmp(op, "down_sqrt")) r = down_sqrt(a);
        else if (!std::strcmp(op, "gamma64")) r = gamma64(static_cast<long long>(a));
        else if (!std::strcmp(op, "norm")) {  // a = max |v|, b = scaled square sum; count on the next token
            long long n = 0;
            if (std::scanf("%lld", &n) != 1) return 2;
            const Interval iv = norm_interval(a, b, 0.0, n);
            r = iv.lower;
            r2 = iv.upper;
            pair = true;
        } else {
            std::printf("unknown %s\n", op);
            return 3;
        }
        if (pair) std::printf("%a %a\n", r, r2);
        else std::printf("%a\n", r);
    }
    return 0;
}
