// Compares amcfd::libm_exp with the C library's exp over seeded random
// arguments in the given range; prints the number of results that differ.
// Usage: exp_check LO HI N [SEED]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../am-cfd_amd/csrc/libm_exp.h"

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const double lo = std::atof(argv[1]), hi = std::atof(argv[2]);
    const long n = std::atol(argv[3]);
    unsigned long long s = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 88172645463325252ull;
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const double x = lo + (hi - lo) * ((s >> 11) * (1.0 / 9007199254740992.0));
        const double a = amcfd::libm_exp(x), b = std::exp(x);
        if (__builtin_bit_cast(unsigned long long, a) != __builtin_bit_cast(unsigned long long, b)) {
            if (bad < 5) std::printf("x=%a mine=%a libm=%a\n", x, a, b);
            ++bad;
        }
    }
    const double special[] = {0.0, -0.0, 1e-300, -1e-300, 0x1p-54, -0x1p-54, 0x1p-55, 709.78, 709.79, -745.1,
                              -745.2, -708.4, -708.3, 512.0, -512.0, 1024.0, -1024.0, 1e308, -1e308,
                              __builtin_inf(), -__builtin_inf(), __builtin_nan("")};
    for (double x : special) {
        const double a = amcfd::libm_exp(x), b = std::exp(x);
        if (__builtin_bit_cast(unsigned long long, a) != __builtin_bit_cast(unsigned long long, b) &&
            !(std::isnan(a) && std::isnan(b))) {
            std::printf("special x=%a mine=%a libm=%a\n", x, a, b);
            ++bad;
        }
    }
    std::printf("mismatches %ld of %ld\n", bad, n);
    return 0;
}
