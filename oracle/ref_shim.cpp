// C entry points over the reference's own src/modulation.cpp, compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`). The
// result, oracle/_ref/libref_modulation.so, is test infrastructure: it pins the
// oracle's and the product's modulation tables bit for bit.
#include <algorithm>
#include <string>
#include "modulation.h"

extern "C" int ref_sps_modulation(double tune, double epsilon, int start, int end, double *out)
{
    std::vector<double> v = sps_modulation(tune, epsilon, start, end);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

extern "C" int ref_basic_modulation(double tune, double omega, double epsilon, int start, int end, double *out)
{
    std::vector<double> v = basic_modulation(tune, omega, epsilon, start, end);
    std::copy(v.begin(), v.end(), out);
    return 0;
}

// pick_a_modulation (modulation.cpp:98-134) for the deterministic kinds
extern "C" int ref_pick_a_modulation(unsigned n_turns, double omega_x, double omega_y, const char *kind,
                                     double omega_0, double epsilon, unsigned offset, double *out_x,
                                     double *out_y)
{
    try {
        auto r = pick_a_modulation(n_turns, omega_x, omega_y, std::string(kind), omega_0, epsilon, offset);
        std::copy(std::get<0>(r).begin(), std::get<0>(r).end(), out_x);
        std::copy(std::get<1>(r).begin(), std::get<1>(r).end(), out_y);
    } catch (...) {
        return 1;
    }
    return 0;
}
