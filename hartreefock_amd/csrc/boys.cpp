// Boys function on the host: F_m(T) = int_0^1 t^{2m} exp(-T t^2) dt.
//
// The reference evaluates F_M(T) from the lower incomplete gamma function and
// recurses downward (HartreeFock/BoysFunction.cpp:7-25, BoysFunctions.cpp:10-25).
// Here F_M(T) comes from the all-positive series
//     F_m(T) = e^{-T} sum_{i>=0} (2T)^i / ((2m+1)(2m+3)...(2m+2i+1))
// in long double (T <= 40), or from F_0 = sqrt(pi/T) erf(sqrt T)/2 and upward
// recursion (T > 40, where the series would need too many terms); lower orders
// follow by the same downward recursion as the reference.
#include <cmath>

#include "hf_internal.h"

namespace hfhost {

namespace {

long double series(int m, long double T) {
    long double term = 1.0L / (2 * m + 1), sum = term;
    for (int i = 1; i < 2000; ++i) {
        term *= 2.0L * T / (2 * m + 2 * i + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
    }
    return std::exp(-T) * sum;
}

}  // namespace

void boys_host(int mmax, double T, double* F) {
    if (T <= 40.0) {
        F[mmax] = (double)series(mmax, T);
        const double e = std::exp(-T);
        for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        const double e = std::exp(-T);
        F[0] = 0.5 * std::sqrt(M_PI / T) * std::erf(std::sqrt(T));
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
    }
}

void boys_table(std::vector<double>& tab) {
    using hfgpu::BOYS_NCOL;
    using hfgpu::BOYS_NROW;
    using hfgpu::BOYS_STRIDE;
    tab.assign((size_t)BOYS_NROW * BOYS_STRIDE, 0.0);
    for (int k = 0; k < BOYS_NROW; ++k) {
        const long double T = (long double)k * 0.1L;
        for (int m = 0; m < BOYS_NCOL; ++m) tab[(size_t)k * BOYS_STRIDE + m] = (double)series(m, T);
        tab[(size_t)k * BOYS_STRIDE + BOYS_NCOL] = (double)std::exp(-T);
    }
}

}  // namespace hfhost

// Boys functions F_0..F_mmax(T) on the host, as IntegralsRepository::getBoysFunctions returns them
// (HartreeFock/IntegralsRepository.cpp:326-342, HartreeFock/BoysFunctions.cpp:10-25).
extern "C" int hf_boys(int mmax, double T, double* F) {
    if (!F || mmax < 0 || mmax > 32 || !(T >= 0) || !std::isfinite(T)) return HF_E_ARG;
    hfhost::boys_host(mmax, T, F);
    return HF_OK;
}
