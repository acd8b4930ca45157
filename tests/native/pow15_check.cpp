// Compares amcfd::pow15 with the C library's pow(v, 1.5) over seeded random
// arguments in [LO, HI]; prints the number of results that differ and the
// largest difference in units in the last place.
// Usage: pow15_check LO HI N [SEED]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../am-cfd_amd/csrc/pow15.h"

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const double lo = std::atof(argv[1]), hi = std::atof(argv[2]);
    const long n = std::atol(argv[3]);
    unsigned long long s = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 88172645463325252ull;
    long bad = 0;
    long long worst = 0;
    for (long i = 0; i < n; ++i) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const double v = lo + (hi - lo) * ((s >> 11) * (1.0 / 9007199254740992.0));
        const double a = amcfd::pow15(v), b = std::pow(v, 1.5);
        const long long ua = __builtin_bit_cast(long long, a), ub = __builtin_bit_cast(long long, b);
        if (ua != ub) {
            if (bad < 5) std::printf("v=%a mine=%a libm=%a\n", v, a, b);
            ++bad;
            const long long d = ua > ub ? ua - ub : ub - ua;
            if (d > worst) worst = d;
        }
    }
    std::printf("mismatches %ld of %ld, max ulp %lld\n", bad, n, worst);
    return 0;
}
