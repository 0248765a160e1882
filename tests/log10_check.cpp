// Host check of hm::log10_fdl (henon_map_cpp_amd/csrc/hm_log10.h) against
// std::log10: the largest error in units in the last place over log-uniform
// arguments across the normal double range, arguments in [0.5, 2] and ratios
// 1 +- 2^-j (j = 1..52), the share of correctly rounded results, and the
// special inputs (zero, negative, infinite, NaN). Prints
// "max_ulp exact_share specials_ok".
#include <cmath>
#include <cstdio>
#include <random>

#include "hm_log10.h"

static double ulp_err(double got, double want)
{
    if (std::isnan(want)) return std::isnan(got) ? 0.0 : 1e300;
    if (got == want) return 0.0;
    const double u = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
    return std::fabs(got - want) / u;
}

int main()
{
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> e(-307.0, 307.0), a(0.5, 2.0), t(-1.0, 1.0);
    double worst = 0.0;
    long exact = 0, total = 0;
    auto check = [&](double x) {
        const double err = ulp_err(hm::log10_fdl(x), std::log10(x));
        worst = std::fmax(worst, err);
        exact += err == 0.0;
        ++total;
    };
    for (int i = 0; i < 2000000; ++i) check(std::pow(10.0, e(g)));
    for (int i = 0; i < 2000000; ++i) check(a(g));
    for (int j = 1; j <= 52; ++j)
        for (int i = 0; i < 20000; ++i) check(1.0 + std::ldexp(t(g), -j));
    const double sp[] = {0.0, -0.0, -1.0, -INFINITY, INFINITY, NAN, 1.0, 10.0, 0.1, 2.2250738585072014e-308,
                         1.7976931348623157e308};
    int ok = 1;
    for (double x : sp) {
        const double gv = hm::log10_fdl(x), wv = std::log10(x);
        if (!(gv == wv || (std::isnan(gv) && std::isnan(wv)))) ok = 0;
    }
    std::printf("%.3f %.4f %d\n", worst, (double)exact / (double)total, ok);
    return 0;
}
